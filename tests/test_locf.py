"""The localized volume force of bla.i loctyp = 3 (locf.f:358-410) on the
spatial Blasius boundary layer of tests/golden/blasius_64x33x8.npz (fringe,
ibc 101): F_i = amp_i exp(-((1 + y)/yscale)^2) f(x) sin(tomeg tc) added to
the nonlinear term.

Pins: the reference bla's run with the force (tests/golden/locf_64x33x8.npz,
amplitudes (0, 0.01, 0.005), x scale 8 at x = -15, y scale 2, frequency 0.3
in delta* units, 20 steps; the force moves w by 7% and v by 1e-3 against
the unforced run): the oracle with the numpy restatement
(oracle/locf_ref.py) and the library on the GPU, after 1 and 20 steps.
Closed forms: the pattern's x profile is the Gaussian's Fourier transform
(exp(-(alpha xscale/2)^2) xscale sqrt(pi)/xl up to the truncation of the
box, for a force centred in it), the y profile exp(-((1 + y)/yscale)^2);
the library's pattern equals the restatement.

Tolerance: velocities 1e-12 relative L2 outside the z-oddball plane (FP64;
2e-5 mixed).
"""
import math
import os

import numpy as np
import pytest

import baseflow_ref
import fringe_ref
import locf_ref
from simson_amd import baseflow, blai, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NX, NY, NZ = 64, 33, 8
DS = 2.0 / 20.0
RE, XL, ZL = 400.0 / DS, 100.0 * DS, 20.0 * DS
DT = 0.1 * DS
FRINGE = (1.0 / DS, -45.0 * DS, 0.0, 15.0 * DS, 10.0 * DS)
AMP = (0.0, 0.01 / DS, 0.005 / DS)
XS, X0, YS, XT, OM = 8.0 * DS, -15.0 * DS, 2.0 * DS, 0.0, 0.3 / DS


@pytest.fixture(scope="module")
def golden():
    return (np.load(os.path.join(ROOT, "tests", "golden", "blasius_64x33x8.npz")),
            np.load(os.path.join(ROOT, "tests", "golden", "locf_64x33x8.npz")))


def rel(a, b):
    a = np.delete(a, NZ // 2, axis=1)
    b = np.delete(b, NZ // 2, axis=1)
    return [float(np.linalg.norm(a[c] - b[c]) / np.linalg.norm(b[c])) for c in range(3)]


def test_pattern_closed_forms_and_restatement():
    P = locf_ref.pattern(NX, NY, NZ, XL, AMP, XS, X0, YS, XT)
    L = gpu.local_force_pattern(NX, NY, XL, *AMP, XS, X0, YS, XT)
    assert np.abs(L - P[:, 0]).max() < 1e-16
    assert np.all(P[:, 1:] == 0)
    # a centred Gaussian: |f_k| = exp(-(alpha xs/2)^2) xs sqrt(pi)/xl
    Pc = gpu.local_force_pattern(NX, NY, XL, 1.0, 0.0, 0.0, XS, 0.0, YS, 0.0)
    eta = np.cos(np.pi * np.arange(NY) / (NY - 1))
    fy = np.exp(-((1 + eta) / YS) ** 2)
    alfa = 2 * np.pi / XL * np.arange(NX // 2)
    fk = np.exp(-(alfa * XS / 2) ** 2) * XS * math.sqrt(math.pi) / XL
    ref = fy[:, None] * fk[None, :] * np.where(np.arange(NX // 2) % 2 == 0, 1.0, -1.0)
    assert np.abs(Pc[0] - ref).max() < 1e-12
    with pytest.raises(gpu.SimsonError):
        gpu.local_force_pattern(NX, NY, XL, 1.0, 0.0, 0.0, XS, 0.0, YS, 2.0)


def test_bla_i_loctyp_3_round_trip():
    p = blai.BlaParams(namnin="a.u", namnut="b.u", tmax=1.0, maxit=3, dt=0.1, spat=True,
                       cflux=False, ibc=101, fmax=1.0, fstart=-4.0, fend=0.0, frise=2.0,
                       ffall=1.0, loctyp=3, lamp=(0.0, 0.01, 0.005), lxscale=8.0,
                       lxloc0=-15.0, lyscale=2.0, lxtype=1.0, ltomeg=0.3)
    q = blai.parse(blai.format_channel(p))
    assert q.loctyp == 3 and q.lamp == (0.0, 0.01, 0.005)
    assert (q.lxscale, q.lxloc0, q.lyscale, q.lxtype, q.ltomeg) == (8.0, -15.0, 2.0, 1.0, 0.3)


def test_oracle_matches_reference(golden):
    from pyoracle import Oracle
    gb, g = golden
    tab = baseflow.fsc_table()
    rb, rbj, rdu = baseflow_ref.base_flow(NX, NY, XL, RE, DS, tab.tab, tab.dstar2, -35 * DS,
                                          25 * DS)
    P = locf_ref.pattern(NX, NY, NZ, XL, AMP, XS, X0, YS, XT)
    o = Oracle(NX, NY, NZ, RE, XL, ZL, cflux=False, retau=0.0)
    try:
        ur, ui, puw = o.new_state(gb["init"])
        o.px = 0.0
        o.set_fringe(fringe_ref.fringe_lambda(3 * NX // 2, XL, *FRINGE), rb)
        o.set_bc(101, rbj, rdu)
        locf_ref.run(o, ur, ui, puw, 1, DT, 0.0, P, OM)
        e1 = rel(ur[:3] + 1j * ui[:3], g["u1"])
        locf_ref.run(o, ur, ui, puw, 19, DT, DT, P, OM)
        e20 = rel(ur[:3] + 1j * ui[:3], g["u20"])
    finally:
        o.close()
    print(f"oracle locf: 1 step {e1}, 20 steps {e20}")
    assert max(e1) <= 1e-12 and max(e20) <= 1e-12, (e1, e20)
    assert max(rel(g["u20"], gb["a_u20"])) > 1e-2          # the force matters


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_gpu_matches_reference(golden, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] loctyp 3 volume force ({form}) {NX}x{NY}x{NZ}", flush=True)
    gb, g = golden
    tab = baseflow.fsc_table()
    base, bj, du0upp = baseflow.spatial_base_flow(NX, NY, XL, RE, DS, tab, -35 * DS, 25 * DS)
    prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
    tol = 1e-12 if form == "fp64" else 2e-5
    with BlaGPU(NX, NY, NZ, RE, XL, ZL, cflux=False, retau=0.0, precision=prec, ibc=101) as G:
        G.fringe(*FRINGE, base=base)
        G.free_stream(bj, du0upp)
        G.local_force(*AMP, xscale=XS, xloc0=X0, yscale=YS, xtype=XT, tomeg=OM)
        G.upload_velocity(gb["init"], 0.0)
        G.run(1, DT, icfl=3)
        e1 = rel(G.velocities(), g["u1"])
        G.run(19, DT, icfl=3, first=False)
        e20 = rel(G.velocities(), g["u20"])
    print(f"    after 1 step {e1}, 20 steps {e20}", flush=True)
    assert max(e1) <= tol and max(e20) <= tol, (e1, e20)
