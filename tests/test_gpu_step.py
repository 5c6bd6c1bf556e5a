"""Parity of the HIP step (through the C ABI) with the oracle and reference bla.

Tolerances, all FP64:
* velocities: <= 1e-12 relative L2 per component against reference bla
  after 100 steps (the FP64 parity bar), and per step against the oracle;
* one half-step from identical inputs: step 2 (H) 1e-13; step 3
  velocities 1e-12: the Helmholtz tau systems for lam = k^2 + 2 Re/(a+b)
  ~ 1e6 have condition numbers ~ ny^2 (a few), so two FP64 codes that round
  differently (FMA contraction on the GPU) differ by ~1e-13 after one solve;
  vorticities and partial right-hand sides 1e-10, because they carry one or
  two Chebyshev derivatives that amplify rounding by up to ny^2;
* px: 1e-9 relative.  The constant-flux update px*(1+(m-m1)/(m2-m1))
  divides a rounding-level flux residual (~1e-15) by the flux response to
  px (~4e-6 at dt = 0.005), so any two correct FP64 implementations that
  round differently (FMA contraction here) differ at ~1e-10.
"""
import math
import sys

import numpy as np
import pytest

import pyoracle
from simson_amd import fields
from simson_amd.gpu import BlaGPU, rk3_weights

pytestmark = pytest.mark.gpu

ODD_MASK = True


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_slots(a, b, nz, skip_oddball=True, nslots=7):
    out = []
    for s in range(nslots):
        x, y = a[s], b[s]
        if skip_oddball:
            x = np.delete(x, nz // 2, axis=0)
            y = np.delete(y, nz // 2, axis=0)
        den = np.linalg.norm(y)
        out.append(float(np.linalg.norm(x - y) / den) if den > 0 else float(np.linalg.norm(x)))
    return out


def c1_field():
    return fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0,
                                kmax=8, rms=1e-3, seed=1)


CASES = [
    # nx, ny, nz, re, xl, zl, kmax, rms, cflux, retau
    (32, 33, 32, 3000.0, 4 * math.pi, 2 * math.pi, 8, 1e-3, True, 0.0),
    (16, 17, 8, 2000.0, 2 * math.pi, math.pi, 3, 1e-2, False, 100.0),
    # FFT lengths 30 (radix 5) take the generic, runtime-sized kernels
    (40, 25, 20, 3000.0, 2 * math.pi, math.pi, 4, 1e-2, True, 0.0),
    (64, 65, 48, 4200.0, 4 * math.pi, 4 * math.pi / 3, 10, 5e-2, True, 0.0),
    (128, 97, 128, 4200.0, 4 * math.pi, 4 * math.pi / 3, 16, 1e-1, True, 0.0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}")
def test_prhs_and_substep_match_oracle(case):
    nx, ny, nz, re, xl, zl, kmax, rms, cflux, retau = case
    say(f"[gpu] prhs+substep parity {nx}x{ny}x{nz}")
    f = fields.channel_field(nx, ny, nz, xl, zl, re, kmax=kmax, rms=rms, seed=7)
    o = pyoracle.Oracle(nx, ny, nz, re, xl, zl, cflux=cflux, retau=retau)
    ur, ui, puw = o.new_state(f.u)
    dt = 0.005
    a, b = [dt * 8 / 15, dt * 5 / 12, dt * 3 / 4], [0., dt * -17 / 60, dt * -5 / 12]
    px0 = o.px
    o.prhs(ur, ui, puw, a[0])
    with BlaGPU(nx, ny, nz, re, xl, zl, cflux=cflux, retau=retau) as g:
        g.upload_velocity(f.u, px0)
        g.prhs(a[0])
        gr, gi, gpuw, gpx = g.download()
        e = rel_slots(gr + 1j * gi, ur + 1j * ui, nz, skip_oddball=False)
        say(f"    prhs rel L2 per slot {['%.1e' % v for v in e]}")
        assert max(e[:3]) == 0.0 and max(e) < 1e-10, e
        assert np.abs(gpuw - puw).max() <= 1e-13 * np.abs(puw).max()
        # step 2 and step 3 each from identical inputs
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
        assert abs(gpx - o.px) <= 1e-9 * abs(o.px)
        gr, gi, gpuw, _ = g.download()
        e = rel_slots(gr + 1j * gi, ur + 1j * ui, nz)
        say(f"    step 3 rel L2 {['%.1e' % v for v in e]}")
        assert max(e[:3]) < 1e-12 and max(e) < 1e-10, e
        assert np.abs(gpuw - puw).max() <= 1e-12 * np.abs(puw).max()
        # the rest of the full step, substep by substep, without re-syncing
        for it in (1, 2):
            cflp = o.nonlin(ur, ui, True)
            o.linear(ur, ui, puw, a[it], b[it], a[(it + 1) % 3], b[(it + 1) % 3])
            gcfl, gpx = g.substep(a[it], b[it], a[(it + 1) % 3], b[(it + 1) % 3], do_cfl=True)
            # states now differ by rounding; CFL only feeds the stop check
            assert abs(gcfl - cflp) <= 1e-12 * cflp
            assert abs(gpx - o.px) <= 1e-9 * abs(o.px)
        gr, gi, gpuw, gpx = g.download()
        e = rel_slots(gr + 1j * gi, ur + 1j * ui, nz)
        say(f"    full step rel L2 {['%.1e' % v for v in e]}")
        assert max(e[:3]) < 1e-12 and max(e) < 1e-10, e
        assert np.abs(gpuw - puw).max() <= 1e-12 * np.abs(puw).max()


def test_c1_matches_reference_bla_100_steps(golden_c1):
    say("[gpu] 32x33x32 100 steps vs reference bla")
    f = c1_field()
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / f.re)
        cfl1 = g.run(1, 0.01, icfl=3)
        u1 = g.velocities()
        px1 = g.download(0)[3]
        g.run(9, 0.01, icfl=3, first=False)
        u10 = g.velocities()
        px10 = g.download(0)[3]
        cfl100 = g.run(90, 0.01, icfl=3, first=False)
        u100 = g.velocities()
    for u, key in ((u1, "u1"), (u10, "u10"), (u100, "u100")):
        ref = golden_c1[key]
        d = np.delete(u - ref, 16, axis=1)
        r = np.delete(ref, 16, axis=1)
        errs = [np.linalg.norm(d[c]) / np.linalg.norm(r[c]) for c in range(3)]
        say(f"    {key}: rel L2 {errs}")
        assert max(errs) <= 1e-12, (key, errs)
        odd = np.abs(u[:, 16] - ref[:, 16]).max() / np.abs(ref[:, 16]).max()
        assert odd < 1e-8
    assert abs(cfl1 - golden_c1["cfl"][0]) <= 1e-12 * golden_c1["cfl"][0]
    assert abs(cfl100 - golden_c1["cfl"][99]) <= 1e-12 * golden_c1["cfl"][99]
    # amp.stat row n holds px after n full steps
    for px, n in ((px1, 1), (px10, 10)):
        assert abs(px - golden_c1["amp_px"][n]) <= 1e-9 * abs(golden_c1["amp_px"][n])


def test_c1_matches_reference_bla_1000_steps():
    """The whole C1 run (1000 steps) against the reference bla's fields after
    500 and 1000 steps, its last CFL*dt line and all 100 amp.stat px rows
    (written after steps 1, 11, ..., 991).  Bars as for the oracle: 1e-12
    after 500 steps; 1e-11 after 1000, since transient growth over steps
    100-1000 amplifies rounding differences ~40x (the oracle ends 2.4e-12
    from the reference in v); px 1e-9."""
    import os
    from conftest import ROOT
    gold = np.load(os.path.join(ROOT, "tests", "golden", "c1_1000_32x33x32.npz"))
    say("[gpu] 32x33x32 1000 steps vs reference bla")
    f = c1_field()
    amp = gold["amp"]
    rows = {int(round(t / 0.01)): i for i, t in enumerate(amp[:, 0])}
    done, px_err, out = 0, 0.0, {}
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / f.re)
        for target in sorted(set(rows) | {500, 1000}):
            cfl = g.run(target - done, 0.01, icfl=3, first=(done == 0))
            done = target
            if target in rows:
                px = g.download(0)[3]
                px_ref = amp[rows[target], 11]
                px_err = max(px_err, abs(px - px_ref) / abs(px_ref))
            if target in (500, 1000):
                out[target] = g.velocities()
    for n, bar in ((500, 1e-12), (1000, 1e-11)):
        ref = gold[f"u{n}"]
        d = np.delete(out[n] - ref, 16, axis=1)
        r = np.delete(ref, 16, axis=1)
        errs = [np.linalg.norm(d[c]) / np.linalg.norm(r[c]) for c in range(3)]
        say(f"    u{n}: rel L2 {errs}")
        assert max(errs) <= bar, (n, errs)
    say(f"    px max rel err over 100 rows {px_err:.2e}")
    assert px_err <= 1e-9
    assert abs(cfl - gold["cfl"][-1]) <= 1e-12 * gold["cfl"][-1]


def test_run_equals_substep_sequence():
    """The fixed-dt driver and host-driven substeps are the same computation
    (weights formed as bla does, dtnp1*anrk, bla.f:2394-2396)."""
    say("[gpu] run() vs host-driven substeps")
    f = c1_field()
    dt = 0.01
    a, b = rk3_weights(dt)
    with BlaGPU.from_ufield(f) as g1, BlaGPU.from_ufield(f) as g2:
        g1.upload_velocity(f.u, -2.0 / f.re)
        g2.upload_velocity(f.u, -2.0 / f.re)
        g1.run(2, dt, icfl=0)
        g2.prhs(a[0])
        for _ in range(2):
            for it in range(3):
                g2.substep(a[it], b[it], a[(it + 1) % 3], b[(it + 1) % 3])
        s1 = g1.download()
        s2 = g2.download()
    for x, y in zip(s1[:3], s2[:3]):
        assert np.array_equal(x, y)
    assert s1[3] == s2[3]


def test_deterministic():
    """Two identical runs give bit-identical states (no atomics, no races)."""
    say("[gpu] determinism")
    f = fields.channel_field(64, 65, 48, 4 * math.pi, 4 * math.pi / 3, 4200.0,
                             kmax=10, rms=5e-2, seed=5)
    out = []
    for _ in range(2):
        with BlaGPU.from_ufield(f) as g:
            g.upload_velocity(f.u, -2.0 / f.re)
            g.run(3, 0.005, icfl=3)
            out.append(g.download())
    for x, y in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(x, y)
    assert out[0][3] == out[1][3]


def test_poiseuille_fixed_point_gpu():
    say("[gpu] laminar Poiseuille fixed point")
    nx, ny, nz = 16, 33, 16
    f = fields.channel_field(nx, ny, nz, 4 * math.pi, 2 * math.pi, 3000.0,
                             kmax=0, rms=0.0, seed=1)
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / f.re)
        g.run(50, 0.01)
        u = g.velocities()
        _, _, _, px = g.download(0)
    assert np.abs(u[0, 0, :, 0] - f.u[0, 0, :, 0]).max() < 1e-13
    rest = u.copy()
    rest[0, 0, :, 0] = 0
    assert np.abs(np.delete(rest, nz // 2, axis=1)).max() < 1e-14
    assert abs(px + 2.0 / 3000.0) < 1e-8 * 2.0 / 3000.0


def test_cfl_stop_code():
    """A field with huge velocities trips the CFL stop of wcflbl.f:93-98."""
    say("[gpu] CFL stop code 324234")
    f = c1_field()
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u * 200.0, -2.0 / f.re)
        with pytest.raises(Exception) as ei:
            g.run(1, 0.5, icfl=3)
        assert getattr(ei.value, "code", None) == 324234


def test_nan_field_stops():
    """A field that holds a NaN fails the CFL check: the CFL maxima take NaN
    as +infinity (fmax would drop it), so the run returns the stop code of
    wcflbl.f:93-98 at the first evaluation instead of running on."""
    say("[gpu] NaN field: CFL stop code")
    f = c1_field()
    u = f.u.copy()
    u[1, 3, 10, 2] = np.nan
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(u, -2.0 / f.re)
        with pytest.raises(Exception) as ei:
            g.run(50, 0.01, icfl=3)
        assert getattr(ei.value, "code", None) == 324234
    # the same run with a finite field is not stopped
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / f.re)
        assert g.run(3, 0.01, icfl=3) < 10.0
