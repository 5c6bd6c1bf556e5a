"""The mixed-precision step (SIMSON_MIXED) against the FP64 step and the oracle.

The mixed form runs the x-z FFTs, the nonlinear products and the y
Chebyshev transforms in FP32 and keeps the tau solves, the mean flow, px
and the partial right-hand sides in FP64.  Its per-step error is therefore
FP32 rounding of H (relative ~1e-7) entering the solution through dt, plus
FP32 rounding of the stored velocities (~6e-8 relative to the largest
component); these tests bound what that gives:
* one substep of step 2 from identical inputs: H within 2e-6 (relative L2);
* the 32x33x32 golden run: after 100 steps every velocity component within
  1e-5 of the reference bla (relative L2 of the perturbation-carrying
  components, whose norms are 1e-3 of the mean flow);
* laminar Poiseuille stays a fixed point to FP32 rounding;
* 128x97x128: mean and rms profiles after 20 steps within 1e-5 of FP64.
"""
import math
import sys

import numpy as np
import pytest

from simson_amd import fields
from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU, SimsonError, rk3_weights

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def c1_field():
    return fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0,
                                kmax=8, rms=1e-3, seed=1)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def profiles(u):
    """Mean U(y) and rms u', v', w' profiles from Fourier coefficients
    u[comp, z, y, kx] (Parseval; kx > 0 modes count twice).  The z-oddball
    plane kz = -nz/2 holds the unnormalised nonlinear term, not velocity
    (both codes zero it before every physical-space use), and is left out."""
    u = np.delete(u, u.shape[1] // 2, axis=1)
    mean = u[:, 0, :, 0].real
    w = np.ones(u.shape[3])
    w[1:] = 2.0
    p = np.abs(u) ** 2 * w[None, None, None, :]
    tot = p.sum(axis=(1, 3)) - np.abs(u[:, 0, :, 0]) ** 2
    return mean, np.sqrt(np.maximum(tot, 0.0))


def test_mixed_rejects_unsupported_sizes():
    # nxp/2 = 30 has no specialised FP32 kernel
    with pytest.raises(SimsonError):
        BlaGPU(40, 25, 20, 3000.0, 2 * math.pi, math.pi, precision=SIMSON_MIXED)


def test_mixed_step2_matches_fp64():
    say("[gpu] mixed step 2 vs FP64")
    f = c1_field()
    a, b = rk3_weights(0.01)
    out = {}
    for prec in (SIMSON_FP64, SIMSON_MIXED):
        with BlaGPU.from_ufield(f, precision=prec) as g:
            g.upload_velocity(f.u, -2.0 / f.re)
            g.prhs(a[0])
            g.nonlin(True)
            out[prec] = g.download(3)
    for s in range(3):
        hm = out[SIMSON_MIXED][0][s] + 1j * out[SIMSON_MIXED][1][s]
        hd = out[SIMSON_FP64][0][s] + 1j * out[SIMSON_FP64][1][s]
        if s == 1:
            # the mixed form leaves the mean x mean part of H2 out of its
            # kx = kz = 0 mode (unused by the v equation there)
            hm[0, :, 0] = 0
            hd[0, :, 0] = 0
        x = np.delete(hm, 16, axis=0)
        y = np.delete(hd, 16, axis=0)
        e = rel(x, y)
        say(f"    H{s + 1} rel L2 {e:.2e}")
        assert e < 2e-6


def test_mixed_c1_100_steps_vs_reference(golden_c1):
    say("[gpu] mixed 32x33x32 100 steps vs reference bla")
    f = c1_field()
    with BlaGPU.from_ufield(f, precision=SIMSON_MIXED) as g:
        g.upload_velocity(f.u, -2.0 / f.re)
        g.run(1, 0.01, icfl=3)
        u1 = g.velocities()
        g.run(99, 0.01, icfl=3, first=False)
        u100 = g.velocities()
    for u, key, tol in ((u1, "u1", 1e-6), (u100, "u100", 1e-5)):
        ref = golden_c1[key]
        d = np.delete(u - ref, 16, axis=1)
        r = np.delete(ref, 16, axis=1)
        errs = [float(np.linalg.norm(d[c]) / np.linalg.norm(r[c])) for c in range(3)]
        say(f"    {key}: rel L2 {['%.2e' % e for e in errs]}")
        assert max(errs) <= tol, (key, errs)


def test_mixed_poiseuille_fixed_point():
    say("[gpu] mixed Poiseuille")
    nx, ny, nz = 32, 33, 32
    f = fields.channel_field(nx, ny, nz, 4 * math.pi, 2 * math.pi, 3000.0, kmax=0,
                             rms=0.0, seed=0)
    with BlaGPU.from_ufield(f, precision=SIMSON_MIXED) as g:
        g.upload_velocity(f.u, -2.0 / f.re)
        g.run(20, 0.01, icfl=0)
        u = g.velocities()
        px = g.download(0)[3]
    y = np.cos(np.pi * np.arange(ny) / (ny - 1))
    err = np.abs(u[0, 0, :, 0].real - (1 - y ** 2)).max()
    pert = np.abs(u).max(axis=(1, 2, 3))
    say(f"    mean-flow error {err:.2e}, |v|max {pert[1]:.2e}, px {px:.12e}")
    assert err < 1e-6
    assert pert[1] < 1e-6 and np.abs(u[2]).max() < 1e-6
    assert abs(px + 2.0 / 3000.0) < 1e-6 * 2.0 / 3000.0


def test_mixed_profiles_128x97x128():
    say("[gpu] mixed vs FP64 profiles at 128x97x128")
    f = fields.channel_field(128, 97, 128, 4 * math.pi, 4 * math.pi / 3, 4200.0, kmax=16,
                             rms=0.1, seed=2)
    res = {}
    for prec in (SIMSON_FP64, SIMSON_MIXED):
        with BlaGPU.from_ufield(f, precision=prec) as g:
            g.upload_velocity(f.u, -2.0 / f.re)
            g.run(20, 0.005, icfl=3)
            res[prec] = g.velocities()
    m64, r64 = profiles(res[SIMSON_FP64])
    m32, r32 = profiles(res[SIMSON_MIXED])
    em = np.abs(m32[0] - m64[0]).max() / np.abs(m64[0]).max()
    er = [float(np.abs(r32[c] - r64[c]).max() / np.abs(r64[c]).max()) for c in range(3)]
    say(f"    mean U rel err {em:.2e}; rms u,v,w rel err {['%.2e' % e for e in er]}")
    assert em < 1e-5 and max(er) < 1e-5
