"""Wall-normal resolutions above 128 points: the parity-split Chebyshev
transforms (ny = 129, 257) and the column solver with two and three
coefficients per lane, through the C ABI.

* One FP64 RK3 step from the oracle's state: velocities 1e-12 relative L2
  as at the other sizes.  The vorticities carry one y derivative and the
  partial right-hand sides up to four (p_v = -(D^2 - lam) phi with
  phi = D^2 v - k^2 v); spectral derivatives amplify rounding by about ny
  each, so their bar grows with ny: measured 3.1e-12 at ny = 129 and
  2.9e-10 at ny = 257 for p_v, bar 1e-10 (129) and 1e-9 (257).  px: the
  constant-flux update divides a rounding-level flux residual by the flux
  response to px, which is proportional to dt (0.002 here, smaller than
  the other tests' 0.005-0.01), measured 8e-9 relative, bar 2e-8.
* The split transforms against the dense ones at ny = 97 (both exist
  there): one step of each, 1e-13.
* The mixed form against FP64 at 32x129x32 and 16x257x16 (mean and rms
  profiles after 20 steps, 1e-5), at the LES grid 128x129x128 and at the
  large channel 384x257x384 (Re_cl 13000, lx 2 pi, lz pi, dt 0.002, 10%
  rms over |kx|, |kz| <= 32).
"""
import math
import os
import sys

import numpy as np
import pytest

import pyoracle
from simson_amd import fields
from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_slots(a, b, nz, nslots=7):
    out = []
    for s in range(nslots):
        x = np.delete(a[s], nz // 2, axis=0)
        y = np.delete(b[s], nz // 2, axis=0)
        den = np.linalg.norm(y)
        out.append(float(np.linalg.norm(x - y) / den) if den > 0 else float(np.linalg.norm(x)))
    return out


def profiles(u):
    """Mean U(y) and rms u', v', w' from Fourier coefficients (Parseval; the
    z-oddball plane is left out)."""
    u = np.delete(u, u.shape[1] // 2, axis=1)
    mean = u[:, 0, :, 0].real
    w = np.ones(u.shape[3])
    w[1:] = 2.0
    p = np.abs(u) ** 2 * w[None, None, None, :]
    tot = p.sum(axis=(1, 3)) - np.abs(u[:, 0, :, 0]) ** 2
    return mean, np.sqrt(np.maximum(tot, 0.0))


CASES = [
    # nx, ny, nz, re, xl, zl, kmax, rms
    (32, 129, 32, 4200.0, 2 * math.pi, math.pi, 8, 5e-2),
    (16, 257, 16, 13000.0, 2 * math.pi, math.pi, 4, 5e-2),
]
PX_BAR = 2e-8


def deriv_bar(ny):
    """bar for the vorticities and partial right-hand sides (slots 3-6)"""
    return 1e-10 if ny <= 129 else 1e-9


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}")
def test_step_matches_oracle(case):
    nx, ny, nz, re, xl, zl, kmax, rms = case
    say(f"[gpu] FP64 step vs oracle {nx}x{ny}x{nz}")
    f = fields.channel_field(nx, ny, nz, xl, zl, re, kmax=kmax, rms=rms, seed=9)
    o = pyoracle.Oracle(nx, ny, nz, re, xl, zl, cflux=True)
    ur, ui, puw = o.new_state(f.u)
    dt = 0.002
    a, b = [dt * 8 / 15, dt * 5 / 12, dt * 3 / 4], [0., dt * -17 / 60, dt * -5 / 12]
    px0 = o.px
    o.prhs(ur, ui, puw, a[0])
    with BlaGPU(nx, ny, nz, re, xl, zl, cflux=True) as g:
        g.upload_velocity(f.u, px0)
        g.prhs(a[0])
        gr, gi, gpuw, _ = g.download()
        e = rel_slots(gr + 1j * gi, ur + 1j * ui, nz)
        say(f"    prhs rel L2 {['%.1e' % v for v in e]}")
        assert max(e[:3]) < 1e-12 and max(e) < deriv_bar(ny), e
        g.upload_state(ur, ui, puw, o.px)
        cflp = o.nonlin(ur, ui, True)
        gcfl = g.nonlin(True)
        assert abs(gcfl - cflp) <= 1e-13 * cflp
        gr, gi, _, _ = g.download(3)
        e = rel_slots(gr + 1j * gi, ur[:3] + 1j * ui[:3], nz, nslots=3)
        say(f"    step 2 rel L2 {['%.1e' % v for v in e]}")
        assert max(e) < 1e-13, e
        g.upload_state(ur, ui, puw, o.px)
        o.linear(ur, ui, puw, a[0], b[0], a[1], b[1])
        gpx = g.linear(a[0], b[0], a[1], b[1])
        assert abs(gpx - o.px) <= PX_BAR * abs(o.px)
        gr, gi, gpuw, _ = g.download()
        e = rel_slots(gr + 1j * gi, ur + 1j * ui, nz)
        say(f"    step 3 rel L2 {['%.1e' % v for v in e]}")
        assert max(e[:3]) < 1e-12 and max(e) < deriv_bar(ny), e
        assert np.abs(gpuw - puw).max() <= 1e-12 * np.abs(puw).max()
        for it in (1, 2):
            o.nonlin(ur, ui, True)
            o.linear(ur, ui, puw, a[it], b[it], a[(it + 1) % 3], b[(it + 1) % 3])
            _, gpx = g.substep(a[it], b[it], a[(it + 1) % 3], b[(it + 1) % 3], do_cfl=True)
            say(f"    substep {it + 1}: px rel err {abs(gpx - o.px) / abs(o.px):.1e}")
            assert abs(gpx - o.px) <= PX_BAR * abs(o.px)
        gr, gi, _, _ = g.download()
        e = rel_slots(gr + 1j * gi, ur + 1j * ui, nz)
        say(f"    full step rel L2 {['%.1e' % v for v in e]}")
        assert max(e[:3]) < 1e-12 and max(e) < deriv_bar(ny), e


def test_split_equals_dense_transforms():
    """At ny = 97 both the split and the dense matrix-core transforms exist;
    SIMSON_CHEB=dense (read at init) keeps the dense ones."""
    say("[gpu] split vs dense transforms 64x97x48, FP64 and mixed")
    f = fields.channel_field(64, 97, 48, 4 * math.pi, 4 * math.pi / 3, 4200.0, kmax=10,
                             rms=5e-2, seed=4)
    for prec, bar in ((SIMSON_FP64, 1e-13), (SIMSON_MIXED, 2e-6)):
        out = []
        for mode in ("dense", None):
            if mode:
                os.environ["SIMSON_CHEB"] = mode
            try:
                with BlaGPU.from_ufield(f, precision=prec) as g:
                    g.upload_velocity(f.u, -2.0 / f.re)
                    g.run(1, 0.005, icfl=3)
                    out.append(g.download())
            finally:
                os.environ.pop("SIMSON_CHEB", None)
        e = rel_slots(out[1][0] + 1j * out[1][1], out[0][0] + 1j * out[0][1], 48)
        say(f"    precision {prec}: rel L2 {['%.1e' % v for v in e]}")
        assert max(e[:3]) < bar and max(e) < 100 * bar, e


def mixed_vs_fp64(nx, ny, nz, re, xl, zl, kmax, rms, dt, nsteps, seed):
    f = fields.channel_field(nx, ny, nz, xl, zl, re, kmax=kmax, rms=rms, seed=seed)
    res = {}
    for prec in (SIMSON_FP64, SIMSON_MIXED):
        with BlaGPU.from_ufield(f, precision=prec) as g:
            g.upload_velocity(f.u, -2.0 / re)
            g.run(nsteps, dt, icfl=3)
            res[prec] = g.velocities()
    m64, r64 = profiles(res[SIMSON_FP64])
    m32, r32 = profiles(res[SIMSON_MIXED])
    em = np.abs(m32[0] - m64[0]).max() / np.abs(m64[0]).max()
    er = [float(np.abs(r32[c] - r64[c]).max() / np.abs(r64[c]).max()) for c in range(3)]
    say(f"    mean U rel err {em:.2e}; rms u,v,w rel err {['%.2e' % e for e in er]}")
    return em, er


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}")
def test_mixed_large_ny_small_grid(case):
    nx, ny, nz, re, xl, zl, kmax, rms = case
    say(f"[gpu] mixed vs FP64 profiles {nx}x{ny}x{nz}, 20 steps")
    em, er = mixed_vs_fp64(nx, ny, nz, re, xl, zl, kmax, rms, 0.002, 20, 3)
    assert em < 1e-5 and max(er) < 1e-5


def test_mixed_les_grid_128x129x128():
    say("[gpu] mixed vs FP64 profiles 128x129x128 (lx 2pi, lz pi), 20 steps")
    em, er = mixed_vs_fp64(128, 129, 128, 4200.0, 2 * math.pi, math.pi, 16, 0.1, 0.004, 20, 6)
    assert em < 1e-5 and max(er) < 1e-5


def test_mixed_large_channel_384x257x384():
    say("[gpu] mixed vs FP64 profiles 384x257x384 (Re_cl 13000), 10 steps")
    em, er = mixed_vs_fp64(384, 257, 384, 13000.0, 2 * math.pi, math.pi, 32, 0.1, 0.002, 10, 5)
    assert em < 1e-5 and max(er) < 1e-5
