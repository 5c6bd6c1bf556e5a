"""The HPF eddy-viscosity model (sgs.i iles = 3): the Smagorinsky model on
the high-pass filtered velocity, with a fixed coefficient or the dynamic
procedure whose B_ij comes from a secondary high-pass filter.

Pins: the reference bla compiled with the eddy-viscosity models
(iiles_smag = 1) on the LES start field, tests/golden/hpf_32x33x32.npz:
cs = 0.15 (cutoff_inv 1.5, iord 5), dynamic with ineg 1 (cutoff_inv 1.5,
iord 5; also saved after 1 step) and dynamic with ineg 0 (cutoff_inv 3,
iord 2), 10 steps at dt = 0.01.  The CPU test steps the oracle with the
numpy restatement (oracle/smag_ref.py with the filters of les_ref) as its
force; the GPU tests run the library with simson_gpu_les.  The model moves
v and w by 1e-4 to 1e-3 relative within 10 steps.
Closed forms of the secondary filter: zero above kx = nx/4 or |kz| = nz/4,
and inside equal to the primary transfer function of the grid with half
the points; the library's secondary transfer function equals the
restatement's.

Tolerance: 1e-12 relative L2 per component outside the z-oddball plane
(FP64); the mixed form 2e-5 against the reference.
"""
import math
import os

import numpy as np
import pytest

import les_ref
import smag_ref
from simson_amd import fields, sgsi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi
CASES = {
    "c": sgsi.SgsParams(iles=3, cutoff_inv=1.5, iord=5, cs=0.15, ineg=1),
    "d1": sgsi.SgsParams(iles=3, cutoff_inv=1.5, iord=5, cs=0.0, ineg=1),
    "d0": sgsi.SgsParams(iles=3, cutoff_inv=3.0, iord=2, cs=0.0, ineg=0),
}


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "hpf_32x33x32.npz"))


def wavenumbers():
    alfa = 2 * math.pi / XL * np.arange(NX // 2)
    kz = np.where(np.arange(NZ) < NZ // 2, np.arange(NZ), np.arange(NZ) - NZ)
    return alfa, 2 * math.pi / ZL * kz.astype(float)


def start():
    return fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=15, rms=1e-2, seed=5)


def rel(a, b):
    a = np.delete(a, NZ // 2, axis=1)
    b = np.delete(b, NZ // 2, axis=1)
    return [float(np.linalg.norm(a[c] - b[c]) / np.linalg.norm(b[c])) for c in range(3)]


def test_secondary_filter_closed_forms():
    for cut, iord in ((1 / 1.5, 5), (1 / 3.0, 2)):
        _, lpf2, _ = smag_ref.hpf_filters(NX, NY, NZ, cut, iord)
        kz = np.where(np.arange(NZ) <= NZ // 2, np.arange(NZ), NZ - np.arange(NZ))
        outside = (np.arange(NX // 2)[None, :] > NX // 4) | (kz[:, None] > NZ // 4)
        assert np.all(lpf2[outside] == 0.0)
        half, _ = les_ref.filter_operators(NX // 2, NY, NZ // 2, cut, iord)
        kept = [z for z in range(NZ) if kz[z] <= NZ // 4]
        for z in kept:
            zh = kz[z] if z <= NZ // 2 else NZ // 2 - kz[z]
            assert np.abs(lpf2[z, :NX // 4] - half[zh, :NX // 4]).max() < 1e-15
        assert abs(lpf2[0, 0] - 1.0) < 1e-15


def test_library_secondary_filter_matches_restatement():
    from simson_amd.gpu import SimsonError, les_filter_secondary
    for cut, iord in ((1 / 1.5, 5), (1 / 3.0, 2)):
        _, lpf2, _ = smag_ref.hpf_filters(NX, NY, NZ, cut, iord)
        assert np.abs(les_filter_secondary(NX, NZ, cut, iord) - lpf2).max() < 1e-14
    with pytest.raises(SimsonError):
        les_filter_secondary(NX, NZ, 1.0, 5)


def test_sgs_params_parse_hpf():
    p = sgsi.parse("20060909\n3\n1.5\n5\n0.\n1\n")
    assert (p.iles, p.cutoff_inv, p.iord, p.cs, p.ineg) == (3, 1.5, 5, 0.0, 1)
    p = sgsi.parse("20060909\n3\n2.0\n3\n0.1\n")
    assert (p.iles, p.cutoff_inv, p.iord, p.cs) == (3, 2.0, 3, 0.1)
    for q in CASES.values():
        assert sgsi.parse(sgsi.format_sgs(q)) == q
    with pytest.raises(sgsi.SgsInputError):
        sgsi.parse("20060909\n3\n1.5\n7\n0.1\n")


@pytest.mark.parametrize("case", ["c", "d1", "d0"])
def test_oracle_matches_reference(golden, case):
    from pyoracle import Oracle
    sp = CASES[case]
    alfa, beta = wavenumbers()
    d2 = smag_ref.delta2(NX, NY, NZ, XL, ZL)
    hp = smag_ref.hpf_filters(NX, NY, NZ, sp.cutoff, sp.iord)
    f = start()
    o = Oracle(NX, NY, NZ, RE, XL, ZL)
    try:
        ur, ui, puw = o.new_state(f.u)
        o.px = -2.0 / RE

        def model(ur, ui):
            tau = smag_ref.les_tau(ur, ui, alfa, beta, NX, d2, sp.cs, sp.ineg, RE, hpf=hp)
            return smag_ref.sgs_force(tau, alfa, beta)

        smag_ref.run(o, ur, ui, puw, 10, 0.01, model)
        errs = rel(ur[:3] + 1j * ui[:3], golden[f"{case}_u10"])
    finally:
        o.close()
    assert max(errs) <= 1e-12, errs


def test_model_effect_is_resolved(golden):
    les = np.load(os.path.join(ROOT, "tests", "golden", "les_32x33x32.npz"))
    smag = np.load(os.path.join(ROOT, "tests", "golden", "smag_32x33x32.npz"))
    for case in CASES:
        eff = rel(golden[f"{case}_u10"], les["dns_u10"])
        assert max(eff) > 1e-5, (case, eff)
    # the fixed-coefficient model on the filtered field differs from the
    # plain one at the same cs
    eff = rel(golden["c_u10"], smag["c_u10"])
    assert max(eff) > 1e-5, eff


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["c", "d1", "d0"])
def test_gpu_matches_reference(golden, case, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] HPF eddy-viscosity LES {case} ({form}) 32x33x32", flush=True)
    f = start()
    prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
    with BlaGPU.from_ufield(f, precision=prec) as g:
        g.les(CASES[case])
        g.upload_velocity(f.u, -2.0 / RE)
        g.run(1, 0.01, icfl=3)
        if case == "d1":
            errs = rel(g.velocities(), golden["d1_u1"])
            print(f"    after 1 step: {errs}", flush=True)
            assert max(errs) <= (1e-12 if form == "fp64" else 2e-5), errs
        g.run(9, 0.01, icfl=3, first=False)
        errs = rel(g.velocities(), golden[f"{case}_u10"])
    print(f"    after 10 steps: {errs}", flush=True)
    assert max(errs) <= (1e-12 if form == "fp64" else 2e-5), errs
