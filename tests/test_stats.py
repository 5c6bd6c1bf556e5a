"""amp.stat statistics: plane sums (boxamp.f) and the wampbl.f line.

Pins: the amp.stat table written by the reference bla in the golden
32x33x32 run.  Row n holds the statistics of the state after n steps, so
rows 1 and 10 are recomputed on the CPU from the golden fields after 1 and
10 steps (vorticities from the velocities by collocation derivatives) with
the oracle's boxamp restatement and the host writer; on the GPU the same
rows come from simson_gpu_amp while the step runs.

Tolerance 1e-9 relative per column: the rms columns are square roots of
sums of ~1e-7 values that the two codes add in different orders, and the
vorticities the CPU path rebuilds from the velocities differ from the
reference's stored ones by collocation rounding (~1e-12 relative); the
pressure-gradient column is compared at 1e-9 as in the step tests.
"""
import math
import sys

import numpy as np
import pytest

import stats_ref
from simson_amd import osmode, stats

NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi


def wavenumbers():
    alfa = 2 * math.pi / XL * np.arange(NX // 2)
    kz = np.where(np.arange(NZ) < NZ // 2, np.arange(NZ), np.arange(NZ) - NZ)
    return alfa, 2 * math.pi / ZL * kz.astype(float)


def vorticities(u, alfa, beta):
    _, D = osmode.cheb_diff(NY)
    du = np.einsum("ij,zjx->zix", D, u[0])
    dw = np.einsum("ij,zjx->zix", D, u[2])
    omx = dw - 1j * beta[:, None, None] * u[1]
    omz = 1j * alfa[None, None, :] * u[1] - du
    return omx, omz


def check_row(got, ref, tag):
    for c in range(1, 19):
        if c == 18:                     # w_b: a rounding-level bulk spanwise velocity
            assert abs(got[c] - ref[c]) < 1e-14, (tag, c, got[c], ref[c])
            continue
        err = abs(got[c] - ref[c]) / max(abs(ref[c]), 1e-300)
        assert err < 1e-9, (tag, stats.AMP_COLUMNS[c], got[c], ref[c], err)


def test_wint_is_clenshaw_curtis():
    w = stats.wint(NY)
    y = np.cos(np.pi * np.arange(NY) / (NY - 1))
    for p in range(0, 12, 2):
        assert abs(np.sum(w * y ** p) - 2.0 / (p + 1)) < 1e-14
    assert abs(np.sum(w * y ** 3)) < 1e-15


def test_fortran_format():
    assert stats.fortran_e(0.01) == "  0.1000000000000000E-001"
    assert stats.fortran_e(-1.25e-4) == " -0.1250000000000000E-003"
    assert stats.fortran_e(2000.0) == "  0.2000000000000000E+004"


@pytest.mark.parametrize("n,key", [(1, "u1"), (10, "u10")])
def test_cpu_rows_match_reference(golden_c1, n, key):
    alfa, beta = wavenumbers()
    u = golden_c1[key]
    omx, omz = vorticities(u, alfa, beta)
    amp = stats_ref.boxamp_ref(u, omx, omz, alfa, beta)
    ref = golden_c1["amp"][n]
    row = stats.amp_row(amp, ref[0], RE, golden_c1["amp_px"][n])
    check_row(row, ref, key)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_gpu_rows_match_reference(golden_c1, form):
    from simson_amd import fields
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] amp.stat rows 0-10 ({form})", flush=True)
    sys.stdout.flush()
    f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=8, rms=1e-3, seed=1)
    ref = golden_c1["amp"]
    with BlaGPU.from_ufield(f, precision=SIMSON_MIXED if form == "mixed" else SIMSON_FP64) as g:
        g.upload_velocity(f.u, -2.0 / RE)
        g.prhs(0.01 * 8.0 / 15.0)             # om_x, om_z of the start field, as bla
        for n in range(11):
            px = g.download(0)[3]
            row = stats.amp_row(g.amp(), 0.01 * n, RE, px)
            if form == "fp64":
                check_row(row, ref[n], n)
            else:
                # FP32 transforms: rms values to ~1e-6 relative
                err = np.abs(row[1:9] - ref[n][1:9]) / np.abs(ref[n][1:9])
                assert err.max() < 1e-5, (n, err)
            g.run(1, 0.01, icfl=0, first=False)
