"""x-y statistics (boxxys.f indices 1-15, wxys.f file layout).

Pins: the xy-statistics file the reference bla wrote for the golden
32x33x32 run of tests/golden/xys_32x33x32.npz (the les_field start, no
model, dt = 0.01, ixys = 9: samples on the fields after 1, 4, 7 and 10 full
steps, sumw = 0.04, written after 12 steps).  The CPU test steps the oracle
through the same schedule and sums the oracle's boxxys restatement; the GPU
tests do the same with simson_gpu_xys, and compare it with the restatement
applied to the GPU's own state.  Closed forms: the z mean of u (index 1)
is the kz = 0 mode transformed in x only, and the z sum of u^2 equals
Parseval's sum over kz.

Tolerance, relative to each index's largest magnitude, against the
reference file: 1e-12 for the statistics of u, v, w and om_y (measured:
<= 6e-14), 1e-10 for those of om_x and om_z (indices 7, 9, 10, 12; measured:
<= 2.6e-11), which carry a wall-normal derivative of the velocity (slots 4-5
of main storage, from the last linear step): the Chebyshev derivative scales
the ~1e-14 velocity differences by about ny^2.  1e-5 for the mixed form
(FP32 transforms).
"""
import math
import os
import sys

import numpy as np
import pytest

import stats_ref
from simson_amd import stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi
DT = 0.01
SAMPLES = (1, 4, 7, 10)        # full steps done before each sample


@pytest.fixture(scope="module")
def golden_xys():
    return np.load(os.path.join(ROOT, "tests", "golden", "xys_32x33x32.npz"))


def wavenumbers():
    alfa = 2 * math.pi / XL * np.arange(NX // 2)
    kz = np.where(np.arange(NZ) < NZ // 2, np.arange(NZ), np.arange(NZ) - NZ)
    return alfa, 2 * math.pi / ZL * kz.astype(float)


def start_field():
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from simson_amd import fields
    return fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=15, rms=1e-2, seed=5)


WALL_DERIV = (6, 8, 9, 11)     # om_x, om_z and their squares (0-based)


def check(got, ref, tol, tag, tol_deriv=None):
    for i in range(ref.shape[0]):
        scale = np.abs(ref[i]).max()
        err = np.abs(got[i] - ref[i]).max() / scale
        t = tol_deriv if (tol_deriv is not None and i in WALL_DERIV) else tol
        assert err < t, (tag, stats.XYS_NAMES[i], err)


def test_boxxys_ref_closed_forms():
    rng = np.random.default_rng(3)
    alfa, beta = wavenumbers()
    ny = 5
    u = rng.standard_normal((3, NZ, ny, NX // 2)) + 1j * rng.standard_normal((3, NZ, ny, NX // 2))
    # the kx = 0 column Hermitian in kz (a real field), the oddball zero
    u[:, NZ // 2] = 0.0
    u[:, 0, :, 0] = u[:, 0, :, 0].real
    for k in range(1, NZ // 2):
        u[:, NZ - k, :, 0] = np.conj(u[:, k, :, 0])
    omx = np.zeros((NZ, ny, NX // 2), complex)
    omz = np.zeros_like(omx)
    out = stats_ref.boxxys_ref(u, omx, omz, alfa, beta, NX, dtn=2.0)
    # z mean of u = x synthesis of the kz = 0 row, times dtn
    x = np.arange(NX)
    mean = np.zeros((ny, NX))
    for k in range(NX // 2):
        c = u[0, 0, :, k]
        mean += (c[:, None] * np.exp(2j * np.pi * k * x / NX)).real * (1 if k == 0 else 2)
    assert np.abs(out[0] - 2.0 * mean).max() < 1e-12 * np.abs(mean).max()
    # x and z sum of v^2 = nx nz sum over the stored modes of |v|^2 with
    # weight 2 for kx > 0 (Parseval)
    wt = np.where(np.arange(NX // 2) > 0, 2.0, 1.0)
    pars = NX * NZ * (np.abs(u[1]) ** 2 * wt).sum(axis=(0, 2))
    got = out[4].sum(axis=1) / 2.0 * NZ
    assert np.abs(got - pars).max() < 1e-11 * pars.max()


def test_xys_file_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    sums = rng.standard_normal((15, 7, 8))
    path = str(tmp_path / "xy.stat")
    stats.write_xys(path, sums, 0.5, RE, XL, ZL, 1.25, 6)
    f = stats.read_xys(path)
    assert (f.nx, f.ny, f.nz, f.fltype) == (8, 7, 6, 1)
    assert f.sumw == 0.5 and f.t == 1.25 and f.re == RE
    assert np.array_equal(f.stats, sums * (1.0 / 0.5))


def test_oracle_matches_reference_file(golden_xys):
    from pyoracle import Oracle
    alfa, beta = wavenumbers()
    f = start_field()
    o = Oracle(NX, NY, NZ, RE, XL, ZL)
    try:
        ur, ui, puw = o.new_state(f.u)
        o.px = -2.0 / RE
        acc = np.zeros((15, NY, NX))
        done = 0
        for n in SAMPLES:
            st, _ = o.run(ur, ui, puw, n - done, DT, icfl=0)
            assert st == 0
            done = n
            u = ur[:3] + 1j * ui[:3]
            acc += stats_ref.boxxys_ref(u, ur[3] + 1j * ui[3], ur[4] + 1j * ui[4], alfa, beta,
                                        NX, DT)
    finally:
        o.close()
    sumw = DT * len(SAMPLES)
    assert abs(float(golden_xys["sumw"]) - sumw) < 1e-15
    check(acc / sumw, golden_xys["xys"], 1e-12, "oracle", tol_deriv=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_gpu_xys_matches_reference_file(golden_xys, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] xy statistics 32x33x32 ({form})", flush=True)
    alfa, beta = wavenumbers()
    f = start_field()
    acc = np.zeros((15, NY, NX))
    prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
    with BlaGPU.from_ufield(f, precision=prec) as g:
        g.upload_velocity(f.u, -2.0 / RE)
        done = 0
        for n in SAMPLES:
            g.run(n - done, DT, icfl=0, first=done == 0)
            done = n
            one = g.xys(DT)
            acc += one
            if form == "fp64":
                ur, ui, _, _ = g.download(5)
                u = ur[:3] + 1j * ui[:3]
                ref = stats_ref.boxxys_ref(u, ur[3] + 1j * ui[3], ur[4] + 1j * ui[4], alfa,
                                           beta, NX, DT)
                check(one, ref, 1e-13, f"restatement after {n} steps")
    if form == "fp64":
        check(acc / (DT * len(SAMPLES)), golden_xys["xys"], 1e-12, form, tol_deriv=1e-10)
    else:
        check(acc / (DT * len(SAMPLES)), golden_xys["xys"], 1e-5, form)


@pytest.mark.gpu
def test_gpu_xys_accumulates_and_rejects_bad_sizes():
    from simson_amd.gpu import BlaGPU
    print("[gpu] xy statistics accumulation", flush=True)
    f = start_field()
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / RE)
        g.prhs(DT * 8.0 / 15.0)
        a = g.xys(DT)
        b = g.xys(DT, out=g.xys(DT))
        assert np.abs(b - 2 * a).max() <= 1e-15 * np.abs(a).max()
        twelve = g.xys(DT, out=np.zeros((12, NY, NX)))
        assert np.array_equal(twelve, a[:12])
        with pytest.raises(ValueError):
            g.xys(DT, out=np.zeros((11, NY, NX)))
        # the C entry point itself refuses nxys outside 12..15
        from simson_amd.gpu import lib
        big = np.zeros((16, NY, NX))
        assert lib().simson_gpu_xys(g.h, DT, 0.0, 0.0, 16, big) == -1
        assert lib().simson_gpu_xys(g.h, DT, 0.0, 0.0, 11, big) == -1
        assert not big.any()
