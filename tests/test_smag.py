"""Eddy-viscosity LES (sgs.i iles = 2): the Smagorinsky model with the
dynamic procedure (Germano/Lilly) or a fixed coefficient.

Pins: the reference bla compiled with the eddy-viscosity models
(iiles_smag = 1) on the LES start field for 10 steps, tests/golden/
smag_32x33x32.npz: dynamic with each clipping (ineg 0, 1, 2) and cs = 0.15.
The CPU test steps the oracle with the numpy restatement of les.f and
sgs_stress.f (oracle/smag_ref.py) as its force; the GPU tests run the
library with simson_gpu_les.  The model moves v and w by 1e-3 to 4e-3
relative within 10 steps, so agreement at 1e-12 pins it to ~1e-9.
Closed forms: the test filter keeps exactly the modes kx < nx/4,
|kz| <= nz/4; the Chebyshev derivative is exact on polynomials.

Tolerance: 1e-12 relative L2 per component outside the z-oddball plane
(FP64); the mixed form 2e-5 against the reference.
"""
import math
import os

import numpy as np
import pytest

import smag_ref
from simson_amd import fields, sgsi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi
CASES = {"d1": (0.0, 1), "d0": (0.0, 0), "d2": (0.0, 2), "c": (0.15, 1)}


@pytest.fixture(scope="module")
def golden_smag():
    return np.load(os.path.join(ROOT, "tests", "golden", "smag_32x33x32.npz"))


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


def test_cutoff_filter_keeps_the_test_band():
    x = np.arange(NX) * XL / NX
    z = np.arange(NZ) * ZL / NZ
    X, Z = np.meshgrid(x, z)
    a1, b1 = 2 * math.pi / XL, 2 * math.pi / ZL
    inside = np.cos((NX // 4 - 1) * a1 * X + (NZ // 4) * b1 * Z)
    outside = np.cos((NX // 4) * a1 * X) + np.sin((NZ // 4 + 1) * b1 * Z)
    p = (inside + outside)[:, None, :]
    f = smag_ref.cutoff_filter(p)
    assert np.abs(f[:, 0] - inside).max() < 1e-13
    assert np.abs(smag_ref.cutoff_filter(f) - f).max() < 1e-13


def test_cheb_dy_exact_on_polynomials():
    y = np.cos(math.pi * np.arange(NY) / (NY - 1))
    f = (3 * y ** 5 - y ** 2 + 0.5)[None, :, None]
    d = smag_ref.cheb_dy(f)[0, :, 0]
    assert np.abs(d - (15 * y ** 4 - 2 * y)).max() < 1e-11


def test_delta2_closed_form():
    d2 = smag_ref.delta2(NX, NY, NZ, XL, ZL)
    y = np.cos(math.pi * np.arange(NY) / (NY - 1))
    j = NY // 2
    dy = math.sqrt((y[j + 1] - y[j]) * (y[j] - y[j - 1]))
    assert abs(d2[j] - (XL / NX * ZL / NZ * dy) ** (2 / 3)) < 1e-15
    assert d2[0] == d2[-1]


@pytest.mark.parametrize("case", ["d1", "d0", "d2", "c"])
def test_oracle_matches_reference(golden_smag, case):
    from pyoracle import Oracle
    cs, ineg = CASES[case]
    alfa, beta = wavenumbers()
    d2 = smag_ref.delta2(NX, NY, NZ, XL, ZL)
    f = start()
    o = Oracle(NX, NY, NZ, RE, XL, ZL)
    try:
        ur, ui, puw = o.new_state(f.u)
        o.px = -2.0 / RE

        def model(ur, ui):
            tau = smag_ref.les_tau(ur, ui, alfa, beta, NX, d2, cs, ineg, RE)
            return smag_ref.sgs_force(tau, alfa, beta)

        smag_ref.run(o, ur, ui, puw, 10, 0.01, model)
        errs = rel(ur[:3] + 1j * ui[:3], golden_smag[f"{case}_u10"])
    finally:
        o.close()
    assert max(errs) <= 1e-12, errs


def test_model_effect_is_resolved(golden_smag):
    les = np.load(os.path.join(ROOT, "tests", "golden", "les_32x33x32.npz"))
    for case in CASES:
        eff = rel(golden_smag[f"{case}_u10"], les["dns_u10"])
        assert max(eff) > 1e-4, (case, eff)


def test_sgs_params_parse_eddy_viscosity():
    p = sgsi.parse("20060909\n2\n0.\n2\n")
    assert (p.iles, p.cs, p.ineg) == (2, 0.0, 2)
    p = sgsi.parse("20060909\n2\n0.1\n")
    assert (p.iles, p.cs, p.ineg) == (2, 0.1, 1)
    assert sgsi.parse(sgsi.format_sgs(p)) == p
    with pytest.raises(sgsi.SgsInputError):
        sgsi.parse("20060909\n2\n0.\n3\n")
    with pytest.raises(sgsi.SgsInputError):
        sgsi.parse("20060909\n5\n")               # iles beyond 3


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["d1", "d0", "d2", "c"])
def test_gpu_matches_reference(golden_smag, case, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] eddy-viscosity LES {case} ({form}) 32x33x32", flush=True)
    cs, ineg = CASES[case]
    f = start()
    prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
    with BlaGPU.from_ufield(f, precision=prec) as g:
        g.les(sgsi.SgsParams(iles=2, cs=cs, ineg=ineg))
        g.upload_velocity(f.u, -2.0 / RE)
        g.run(1, 0.01, icfl=3)
        if case == "d1":
            errs = rel(g.velocities(), golden_smag["d1_u1"])
            print(f"    after 1 step: {errs}", flush=True)
            assert max(errs) <= (1e-12 if form == "fp64" else 2e-5), errs
        g.run(9, 0.01, icfl=3, first=False)
        errs = rel(g.velocities(), golden_smag[f"{case}_u10"])
    print(f"    after 10 steps: {errs}", flush=True)
    assert max(errs) <= (1e-12 if form == "fp64" else 2e-5), errs
