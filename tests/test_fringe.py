"""Fringe region of spatial simulations (bla.i spat: fring.f, fringp.f,
ffun.f), on the spatial Poiseuille channel (fltype 4, ibc 0, no pressure
gradient).

Pins: the reference bla's spatial-channel runs (tests/golden/
fringe_32x33x32.npz, two fringe settings, the LES start field, 10 steps):
the oracle stepped with the numpy restatement of the fringe function
(oracle/fringe_ref.py) as its forcing, and the library on the GPU.  The
fringe moves v and w by about 4% within 10 steps, so agreement at 1e-12
pins it.  Closed forms: the smooth step of step.f (1/2 at x = 1/2, 0 and 1
outside (0.02, 0.98]); lambda is fmax on the plateau between the rise and
the fall and zero away from the fringe; the library's fringe function
equals the restatement.

Tolerance: velocities 1e-12 relative L2 outside the z-oddball plane (FP64;
2e-5 mixed); pressure 1e-9 as in tests/test_pressure.py.
"""
import dataclasses
import math
import os

import numpy as np
import pytest

import fringe_ref
from simson_amd import blai, fields

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi
NXP = 3 * NX // 2
CASES = {"a": (1.0, -4.0, 0.0, 2.0, 1.0), "b": (2.5, -6.0, 3.0, 3.0, 2.0)}


@pytest.fixture(scope="module")
def golden_fringe():
    return np.load(os.path.join(ROOT, "tests", "golden", "fringe_32x33x32.npz"))


def start():
    f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=15, rms=1e-2, seed=5)
    return dataclasses.replace(f, fltype=4, extra={"bstart": 0.0, "blength": 0.0})


def rel(a, b):
    a = np.delete(a, NZ // 2, axis=1)
    b = np.delete(b, NZ // 2, axis=1)
    return [float(np.linalg.norm(a[c] - b[c]) / np.linalg.norm(b[c])) for c in range(3)]


def prel(p, ref):
    d = np.delete(p - ref, NZ // 2, axis=0)
    return float(np.linalg.norm(d) / np.linalg.norm(np.delete(ref, NZ // 2, axis=0)))


def test_step_closed_forms():
    assert fringe_ref.step(0.5) == 0.5
    assert fringe_ref.step(0.02) == 0.0 and fringe_ref.step(-3.0) == 0.0
    assert fringe_ref.step(0.99) == 1.0
    xs = np.linspace(0.03, 0.97, 50)
    v = [fringe_ref.step(x) for x in xs]
    assert all(b > a for a, b in zip(v, v[1:]))
    assert abs(fringe_ref.step(0.3) + fringe_ref.step(0.7) - 1.0) < 1e-15   # odd symmetry


def test_lambda_plateau_and_support():
    fmax, fstart, fend, frise, ffall = CASES["b"]
    lam = fringe_ref.fringe_lambda(NXP, XL, fmax, fstart, fend, frise, ffall)
    x = -XL / 2 + np.arange(NXP) * XL / NXP
    plateau = (x > fstart + frise) & (x < fend - ffall)
    assert plateau.any() and np.all(lam[plateau] == fmax)
    assert np.all(lam[(x < fstart) | (x > fend)] == 0.0)
    assert np.all(lam >= 0.0) and lam.max() == fmax


def test_library_fringe_function_matches_restatement():
    from simson_amd.gpu import SimsonError, fringe_function
    for fmax, fstart, fend, frise, ffall in CASES.values():
        a = fringe_function(NX, XL, fmax, fstart, fend, frise, ffall)
        b = fringe_ref.fringe_lambda(NXP, XL, fmax, fstart, fend, frise, ffall)
        assert np.abs(a - b).max() < 1e-15
    with pytest.raises(SimsonError):
        fringe_function(NX, XL, 1.0, 0.0, -1.0, 2.0, 1.0)       # fstart >= fend


def test_bla_i_spatial_round_trip():
    p = blai.BlaParams(namnin="a.u", namnut="b.u", dt=0.01, nst=3, xl=XL, cflux=False,
                       retau=0.0, spat=True, fmax=1.0, fstart=-4.0, fend=0.0, frise=2.0,
                       ffall=1.0)
    q = blai.parse(blai.format_channel(p))
    assert q.spat and (q.fmax, q.fstart, q.fend, q.frise, q.ffall) == (1.0, -4.0, 0.0, 2.0, 1.0)
    with pytest.raises(blai.BlaInputError):
        blai.parse(blai.format_channel(dataclasses.replace(p, fstart=1.0)))


@pytest.mark.parametrize("case", ["a", "b"])
def test_oracle_matches_reference(golden_fringe, case):
    from pyoracle import Oracle
    fmax, fstart, fend, frise, ffall = CASES[case]
    f = start()
    o = Oracle(NX, NY, NZ, RE, XL, ZL, cflux=False, retau=0.0)
    try:
        ur, ui, puw = o.new_state(f.u)
        o.px = 0.0
        o.set_fringe(fringe_ref.fringe_lambda(NXP, XL, fmax, fstart, fend, frise, ffall),
                     fringe_ref.poiseuille_base(NY, NXP))
        st, _ = o.run(ur, ui, puw, 1, 0.01)
        assert st == 0
        e1 = rel(ur[:3] + 1j * ui[:3], golden_fringe[f"{case}_u1"])
        st, _ = o.run(ur, ui, puw, 9, 0.01)
        e10 = rel(ur[:3] + 1j * ui[:3], golden_fringe[f"{case}_u10"])
        ep = prel(o.pressure(ur, ui), golden_fringe[f"{case}_p10"])
    finally:
        o.close()
    print(f"oracle fringe {case}: {e1} {e10} pressure {ep:.2e}")
    assert max(e1) <= 1e-12 and max(e10) <= 1e-12, (e1, e10)
    assert ep < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["a", "b"])
def test_gpu_matches_reference(golden_fringe, case, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] fringe {case} ({form}) 32x33x32", flush=True)
    fmax, fstart, fend, frise, ffall = CASES[case]
    f = start()
    prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
    tol = 1e-12 if form == "fp64" else 2e-5
    with BlaGPU(NX, NY, NZ, RE, XL, ZL, cflux=False, retau=0.0, precision=prec) as g:
        g.fringe(fmax, fstart, fend, frise, ffall, fringe_ref.poiseuille_base(NY, NXP))
        g.upload_velocity(f.u, 0.0)
        g.run(1, 0.01, icfl=3)
        e1 = rel(g.velocities(), golden_fringe[f"{case}_u1"])
        g.run(9, 0.01, icfl=3, first=False)
        e10 = rel(g.velocities(), golden_fringe[f"{case}_u10"])
        ep = prel(g.pressure(), golden_fringe[f"{case}_p10"])
    print(f"    after 1 step {e1}, 10 steps {e10}, pressure {ep:.2e}", flush=True)
    assert max(e1) <= tol and max(e10) <= tol, (e1, e10)
    assert ep < (1e-9 if form == "fp64" else 1e-5), ep
