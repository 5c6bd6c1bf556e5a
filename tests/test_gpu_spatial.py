"""The spatial boundary layer at the width of the Tollmien-Schlichting
configuration (nx = 512: x rows of 768 points, transformed as 384-point
complex rows), through the C ABI.

* One FP64 RK3 step of a 512x33x8 Blasius boundary layer (fringe, ibc 101,
  lower-wall strip at F = 100, noise) against the oracle with the same base
  flow: velocities 1e-12 relative L2 as at the other sizes (measured
  4.4e-13).  This pins the 384-point x rows in the FP64 form; the mixed form
  is compared with FP64 after 10 steps at 2e-5 (the bar of the 64x33x8
  reference comparison; measured 1.3e-7).
* The full configuration, 512x97x64 at Re_delta* = 400 (box 1000 x 30 x 50
  delta*, fringe on [-100, 0], strip amplitude 1e-4 on [40, 60] at F = 100,
  noise rms 1e-6, dt = 0.02): 5000 steps in FP64 and in the mixed form.  The
  wave the strip launches is measured by its amplitude along x, the largest
  |u - U| over y in the kz = 0 plane; the mixed run's amplitude profile
  over the 300 delta* downstream of the strip must stay within 2% of
  FP64's there, and the wave must stand 100 times above the noise.
  Measured: amplitude 1.2e-3 next to the strip, 2.3e-5 300 delta*
  downstream after 100 time units; mixed vs FP64 3.4e-5 of the maximum.
"""
import sys

import numpy as np
import pytest

import fringe_ref
import wall_ref
from simson_amd import baseflow
from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b, nz):
    a = np.delete(a, nz // 2, axis=1)
    b = np.delete(b, nz // 2, axis=1)
    return [float(np.linalg.norm(a[c] - b[c]) / np.linalg.norm(b[c])) for c in range(3)]


def run_gpu(case, nx, ny, nz, prec, nsteps, xl, zl):
    f = case.field
    with BlaGPU(nx, ny, nz, f.re, xl, zl, cflux=False, retau=0.0, precision=prec,
                ibc=101) as g:
        g.fringe(*case.fringe, base=case.base)
        g.free_stream(case.bj, case.du0upp)
        if case.wall is not None:
            g.wall(*case.wall)
        g.upload_velocity(f.u, 0.0)
        g.run(nsteps, case.dt, icfl=3)
        return g.velocities()


def test_x384_step_matches_oracle():
    from pyoracle import Oracle
    nx, ny, nz = 512, 33, 8
    say(f"[gpu] spatial Blasius {nx}x{ny}x{nz}: one FP64 step vs oracle")
    c = baseflow.tollmien_schlichting_case(nx, ny, nz, noise=1e-3, kmax=6, seed=7,
                                           strip=(1e-3, 1.0, 40.0, 60.0, 5.0, 5.0, 0.1))
    f = c.field
    o = Oracle(nx, ny, nz, f.re, f.xl, f.zl, cflux=False, retau=0.0)
    try:
        ur, ui, puw = o.new_state(f.u)
        o.px = 0.0
        o.set_fringe(fringe_ref.fringe_lambda(3 * nx // 2, f.xl, *c.fringe), c.base)
        o.set_bc(101, c.bj, c.du0upp)
        wall_ref.run(o, ur, ui, puw, 1, c.dt, 0.0,
                     lambda tc: wall_ref.wall_v(nx, nz, f.xl, f.zl, *c.wall, tc))
        ref = ur[:3] + 1j * ui[:3]
    finally:
        o.close()
    u64 = run_gpu(c, nx, ny, nz, SIMSON_FP64, 1, f.xl, f.zl)
    e = rel(u64, ref, nz)
    say(f"    FP64 vs oracle after 1 step: {['%.1e' % v for v in e]}")
    assert max(e) < 1e-12, e
    a = run_gpu(c, nx, ny, nz, SIMSON_FP64, 10, f.xl, f.zl)
    b = run_gpu(c, nx, ny, nz, SIMSON_MIXED, 10, f.xl, f.zl)
    e = rel(b, a, nz)
    say(f"    mixed vs FP64 after 10 steps: {['%.1e' % v for v in e]}")
    assert max(e) < 2e-5, e


def wave_amplitude(u, base, nx):
    """max over y of |u - U| in the kz = 0 plane at the nx physical x points."""
    nxp = 3 * nx // 2
    bh = np.fft.fft(base[0], axis=1)[:, :nx // 2] / nxp
    d = u[0, 0] - bh                                   # [y][kx]
    full = np.zeros((d.shape[0], nx), complex)
    full[:, :nx // 2] = d
    full[:, nx // 2 + 1:] = np.conj(d[:, 1:][:, ::-1])
    phys = np.fft.ifft(full, axis=1).real * nx
    return np.abs(phys).max(axis=0)


def test_ts_wave_mixed_matches_fp64():
    nx, ny, nz = 512, 97, 64
    nsteps = 5000
    say(f"[gpu] Tollmien-Schlichting case {nx}x{ny}x{nz}, {nsteps} steps, FP64 and mixed")
    c = baseflow.tollmien_schlichting_case()
    f = c.field
    amps = {}
    for name, prec in (("fp64", SIMSON_FP64), ("mixed", SIMSON_MIXED)):
        u = run_gpu(c, nx, ny, nz, prec, nsteps, f.xl, f.zl)
        amps[name] = wave_amplitude(u, c.base, nx)
        say(f"    {name}: done")
    x = (np.arange(nx) / nx - 0.5) * 1000.0             # delta* of the inflow
    win = (x >= 40.0) & (x <= 340.0)
    a64, am = amps["fp64"][win], amps["mixed"][win]
    dev = float(np.abs(am - a64).max() / a64.max())
    say(f"    wave amplitude max {a64.max():.3e} at x = {x[win][np.argmax(a64)]:.0f}; "
        f"mixed vs FP64 {dev:.2e}")
    say("    amplitude along x (FP64): " +
        " ".join(f"{xx:.0f}:{aa:.1e}" for xx, aa in zip(x[win][::16], a64[::16])))
    assert a64.max() > 100 * 1e-6
    assert dev < 2e-2, dev


@pytest.mark.parametrize("nx", [384, 512])
def test_long_x_rows_match_oracle(nx):
    """Channel steps with x rows of 576 and 768 points (the three-pass
    split form of the x-row kernel, nx/2 = 288 and 384 complex points),
    FP64 against the oracle after two steps (1e-12), and the pressure of the
    same state (its x-row variant, 1e-9 as in tests/test_pressure.py)."""
    from pyoracle import Oracle
    from simson_amd import fields
    ny, nz, re = 33, 8, 3000.0
    xl, zl = 4 * np.pi, np.pi
    say(f"[gpu] channel {nx}x{ny}x{nz}: two FP64 steps and the pressure vs oracle")
    f = fields.channel_field(nx, ny, nz, xl, zl, re, kmax=12, rms=5e-2, seed=11)
    o = Oracle(nx, ny, nz, re, xl, zl)
    try:
        ur, ui, puw = o.new_state(f.u)
        o.run(ur, ui, puw, 2, 0.005, icfl=0)
        ref = ur[:3] + 1j * ui[:3]
        pref = o.pressure(ur, ui)
    finally:
        o.close()
    with BlaGPU(nx, ny, nz, re, xl, zl, cflux=True) as g:
        g.upload_velocity(f.u, -2.0 / re)
        g.run(2, 0.005, icfl=0)
        u = g.velocities()
        p = g.pressure()
    e = rel(u, ref, nz)
    d = np.delete(p - pref, nz // 2, axis=0)
    ep = float(np.linalg.norm(d) / np.linalg.norm(np.delete(pref, nz // 2, axis=0)))
    say(f"    velocities {['%.1e' % v for v in e]}, pressure {ep:.1e}")
    assert max(e) < 1e-12, e
    assert ep < 1e-9, ep
