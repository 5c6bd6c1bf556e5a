"""Blowing and suction at the lower wall (bla.i wbci = 1: cwallbc.f and the
inhomogeneous v boundary condition of linearbl.f:458-469).

Pins: the reference bla's temporal channel with the wall strip (tests/
golden/wall_32x33x32.npz: an oscillating strip whose suction part is
damped, and a steady one; the LES start field, 10 steps): the oracle
stepped with the numpy restatement of the wall velocity (oracle/wall_ref.py)
and the library on the GPU, including the CFL*dt log lines (boxcfl.f
without the wall cap).  The strip changes v and w at O(1) within 10 steps.
Closed forms: the reference's own v at the lower wall after a step equals
the restated wall velocity at the start of the last substep; the imposed v
is the wall value exactly; the library's host-side wall spectra equal the
restatement for both signs of sin(tomeg t).

Tolerance: velocities 1e-12 relative L2 outside the z-oddball plane (FP64;
2e-5 mixed); CFL 1e-10.
"""
import math
import os

import numpy as np
import pytest

import wall_ref
from simson_amd import blai, fields

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi
DT = 0.01
CASES = {"a": (0.05, 0.5, -3.0, 2.0, 1.0, 1.5, 2.0, 20.0),
         "b": (0.02, 1.0, -5.0, -1.0, 2.0, 0.5, 1.0, 0.0)}


@pytest.fixture(scope="module")
def golden_wall():
    return np.load(os.path.join(ROOT, "tests", "golden", "wall_32x33x32.npz"))


def start():
    return fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=15, rms=1e-2, seed=5)


def rel(a, b):
    a = np.delete(a, NZ // 2, axis=1)
    b = np.delete(b, NZ // 2, axis=1)
    return [float(np.linalg.norm(a[c] - b[c]) / np.linalg.norm(b[c])) for c in range(3)]


def wall_of(case):
    p = CASES[case]
    return lambda tc: wall_ref.wall_v(NX, NZ, XL, ZL, *p, tc)


def test_reference_wall_value_is_the_restated_strip(golden_wall):
    for case in CASES:
        v = golden_wall[f"{case}_u1"][1][:, -1, :].copy()       # y = -1
        w = wall_of(case)(DT * 2.0 / 3.0)                        # last substep start
        v[NZ // 2] = 0.0
        v[0, 0] = w[0, 0].real
        assert np.abs(v - w).max() < 1e-16, case


def test_library_wall_velocity_matches_restatement():
    from simson_amd.gpu import wall_velocity
    for case, p in CASES.items():
        for tc in (0.003, 0.2, 1.0):
            a = wall_velocity(NX, NZ, XL, ZL, tc, *p)
            b = wall_of(case)(tc)
            assert np.abs(a - b).max() <= 1e-15 * max(1e-300, np.abs(b).max()) + 1e-20


def test_bla_i_wbci_round_trip():
    p = blai.BlaParams(namnin="a.u", namnut="b.u", dt=0.01, nst=3, xl=XL, wbci=1, wamp=0.1,
                       wdamp=0.3, wxstart=-1.0, wxend=1.0, wxrise=0.5, wxfall=0.5, wzbet=2.0,
                       wtomeg=3.0)
    q = blai.parse(blai.format_channel(p))
    assert (q.wbci, q.wamp, q.wdamp, q.wzbet, q.wtomeg) == (1, 0.1, 0.3, 2.0, 3.0)


@pytest.mark.parametrize("case", ["a", "b"])
def test_oracle_matches_reference(golden_wall, case):
    from pyoracle import Oracle
    f = start()
    o = Oracle(NX, NY, NZ, RE, XL, ZL)
    try:
        ur, ui, puw = o.new_state(f.u)
        o.px = -2.0 / RE
        wall_ref.run(o, ur, ui, puw, 10, DT, 0.0, wall_of(case))
        e = rel(ur[:3] + 1j * ui[:3], golden_wall[f"{case}_u10"])
    finally:
        o.close()
    print(f"oracle wall {case}: {e}")
    assert max(e) <= 1e-12, e


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["a", "b"])
def test_gpu_matches_reference(golden_wall, case, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] wall strip {case} ({form}) 32x33x32", flush=True)
    f = start()
    prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
    tol = 1e-12 if form == "fp64" else 2e-5
    with BlaGPU.from_ufield(f, precision=prec) as g:
        g.wall(*CASES[case])
        g.set_time(0.0)
        g.upload_velocity(f.u, -2.0 / RE)
        g.run(1, DT, icfl=3)
        e1 = rel(g.velocities(), golden_wall[f"{case}_u1"])
        cfls = []
        for _ in range(9):
            cfls.append(g.run(1, DT, icfl=3, first=False))
        e10 = rel(g.velocities(), golden_wall[f"{case}_u10"])
    print(f"    after 1 step {e1}, 10 steps {e10}", flush=True)
    assert max(e1) <= tol and max(e10) <= tol, (e1, e10)
    if form == "fp64":
        # run(1) reports the CFL of its first substep: the log lines of steps 2-10
        ref = golden_wall[f"{case}_cfl"][1:]
        np.testing.assert_allclose(cfls, ref, rtol=1e-10)
