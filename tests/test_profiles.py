"""Host-side statistics of simson_amd.profiles and the byte counts of
bench.py, against closed forms (CPU only)."""
import math

import numpy as np

import bench
from simson_amd.profiles import deviation, profiles, wall_shear, wall_slopes


def test_wall_slopes_of_poiseuille():
    """Exact for polynomials up to rounding, which the k^2 weights of the
    wall derivative amplify (~1e-10 at ny = 257)."""
    for ny in (17, 33, 97, 257):
        y = np.cos(np.pi * np.arange(ny) / (ny - 1))
        up, lo = wall_slopes(1.0 - y ** 2)
        assert abs(up + 2.0) < 1e-9 and abs(lo - 2.0) < 1e-9
        # a cubic: U = y^3 has U'(+-1) = 3
        up, lo = wall_slopes(y ** 3)
        assert abs(up - 3.0) < 1e-9 and abs(lo - 3.0) < 1e-9
    y = np.cos(np.pi * np.arange(97) / 96)
    assert abs(wall_shear(1.0 - y ** 2, 4200.0) - 2.0 / 4200.0) < 1e-14


def test_profiles_of_single_modes():
    """u = U(y) + a cos(alpha x) f(y), v = b cos(alpha x) g(y): rms by
    Parseval (a^2 f^2 / 2), uv = a b f g / 2."""
    nz, ny, nxh = 8, 17, 4
    y = np.cos(np.pi * np.arange(ny) / (ny - 1))
    u = np.zeros((3, nz, ny, nxh), complex)
    u[0, 0, :, 0] = 1.0 - y ** 2
    f, g = 1.0 - y ** 2, y * (1.0 - y ** 2)
    a, b = 0.3, 0.2
    # cos(alpha x) = (e^{i alpha x} + e^{-i alpha x}) / 2: coefficient 1/2 at kx = 1
    u[0, 0, :, 1] = a * f / 2
    u[1, 0, :, 1] = b * g / 2
    # the z-oddball plane holds no velocity and must be ignored
    u[:, nz // 2] = 1e3
    mean, rms, uv = profiles(u)
    assert np.allclose(mean[0], 1.0 - y ** 2)
    assert np.allclose(rms[0], np.abs(a * f) / math.sqrt(2))
    assert np.allclose(rms[1], np.abs(b * g) / math.sqrt(2))
    assert np.allclose(rms[2], 0.0)
    assert np.allclose(uv, a * b * f * g / 2)
    assert deviation(mean[0] * (1 + 1e-6), mean[0]) <= 1.0001e-6


def test_bench_step_bytes_follow_the_reference_data_flow():
    """FP64 bytes per RK3 step = 3 substeps x 20 fields x F x 16 B: 762.8 MB
    at 128x97x128 and 18.19 GB at 384x257x384; the mixed form stores 16 of
    the 20 fields in FP32."""
    F2 = 128 * 97 * 64
    assert bench.step_bytes(False) == 3 * 20 * F2 * 16
    assert abs(bench.step_bytes(False) / 1e6 - 762.8) < 0.1
    assert abs(bench.step_bytes(False, bench.LARGE) / 1e9 - 18.19) < 0.01
    assert bench.step_bytes(True) == 3 * (16 * F2 * 8 + 4 * F2 * 16)
    alg = bench.alg_per_launch(True)
    F3 = 127 * 97 * 64
    assert alg["tau"] == ("hbm", 8 * F3 * 8 + 4 * F3 * 16)
    assert alg["zsyn"][1] + alg["zana"][1] == bench.step2_bytes(True)
    # parity-split transform flops: 2 ((H+1)^2 + H^2) per real column
    assert alg["cheb_fwd"][1] == 2.0 * (49 ** 2 + 48 ** 2) * 128 * 127 * 3
