"""Pressure (nonlinp.f + linearp.f, the slot bla writes to <name>.p).

Pins:
* the Neumann tau solver of the oracle against the equations themselves
  (interior rows of u'' - lam u = f and the wall derivatives), the cosh
  solution, and the lam = 0 case whose constant is set to zero;
* the oracle's whole pressure against the t1.u.p / t10.u.p files the
  reference bla (built with pressure = 1) wrote for the golden 32x33x32 run
  (tests/golden/p_32x33x32.npz);
* on the GPU, simson_gpu_pressure against the same files and against the
  oracle at 64x65x48.

Tolerance 1e-9 relative L2 over the planes other than the z oddball (which
holds the transform of u.u, compared at 1e-9 of its own size): the mean
pressure comes from H2 of the mean flow, an O(1e-5) balance of O(1) terms,
so two FP64 codes agree there to ~1e-11 absolute; measured 5e-11 for the
oracle.  The linearp.f wall term -(v - v_prev)/dt is left out (v = 0 at the
walls to rounding, ~1e-18 here).
"""
import math
import sys

import numpy as np
import pytest

import pyoracle
from simson_amd import fields

NX, NY, NZ, RE = 32, 33, 32, 3000.0
XL, ZL = 4 * math.pi, 2 * math.pi


@pytest.fixture(scope="module")
def golden_p():
    from conftest import ROOT
    import os
    return np.load(os.path.join(ROOT, "tests", "golden", "p_32x33x32.npz"))


def cheb_matrix(n):
    y = np.cos(np.pi * np.arange(n) / (n - 1))
    return y, np.cos(np.outer(np.arccos(y), np.arange(n)))     # T_k(y_j)


def test_neumann_solver_satisfies_the_equations():
    L = pyoracle.lib()
    n = 33
    rng = np.random.default_rng(5)
    for lam in (0.5, 30.0, 1.0e4):
        f = rng.standard_normal(n)
        u = np.zeros(n)
        L.orc_tau_solve_neumann(n, lam, f, 0.3, -0.7, u)
        # interior: coefficients of u'' - lam u equal f up to degree n-3
        d2 = np.polynomial.chebyshev.chebder(u, 2)
        res = np.zeros(n)
        res[:len(d2)] = d2
        res -= lam * u
        assert np.abs(res[:n - 2] - f[:n - 2]).max() < 1e-9 * max(1.0, np.abs(f).max() * lam)
        # wall derivatives: u'(1) = sum k^2 u_k, u'(-1) = sum (-1)^(k+1) k^2 u_k
        k2 = np.arange(n) ** 2.0
        up, um = np.sum(k2 * u), np.sum((-1.0) ** (np.arange(n) + 1) * k2 * u)
        assert abs(0.5 * (up - um) - 0.3) < 1e-10 and abs(0.5 * (up + um) + 0.7) < 1e-10


def test_neumann_cosh_and_singular_constant():
    L = pyoracle.lib()
    n = 33
    y, T = cheb_matrix(n)
    lam = 4.0
    # u = cosh(2y): u'' - 4u = 0, u'(+-1) = +-2 sinh 2
    u = np.zeros(n)
    L.orc_tau_solve_neumann(n, lam, np.zeros(n), 2 * math.sinh(2.0), 0.0, u)
    assert np.abs(T @ u - np.cosh(2 * y)).max() < 1e-12
    # lam = 0: u'' = 2 with u'(+-1) = +-2 -> u = y^2 + const, constant set to 0
    f = np.zeros(n)
    f[0] = 2.0
    L.orc_tau_solve_neumann(n, 0.0, f, 2.0, 0.0, u)
    assert abs(u[0]) < 1e-14 and abs(u[2] - 0.5) < 1e-13


def masked_rel(p, ref):
    d = np.delete(p - ref, NZ // 2, axis=0)
    r = np.delete(ref, NZ // 2, axis=0)
    odd = np.abs(p[NZ // 2] - ref[NZ // 2]).max() / np.abs(ref[NZ // 2]).max()
    return float(np.linalg.norm(d) / np.linalg.norm(r)), float(odd)


def test_oracle_pressure_matches_reference(golden_p):
    f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=8, rms=1e-3, seed=1)
    o = pyoracle.Oracle(NX, NY, NZ, RE, XL, ZL)
    ur, ui, puw = o.new_state(f.u)
    o.run(ur, ui, puw, 1, 0.01, icfl=0)
    e, odd = masked_rel(o.pressure(ur, ui), golden_p["p1"])
    print(f"oracle pressure after 1 step: rel L2 {e:.2e}, oddball {odd:.2e}")
    assert e < 1e-9 and odd < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_gpu_pressure_matches_reference(golden_p, form):
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    print(f"[gpu] pressure vs reference ({form})", flush=True)
    sys.stdout.flush()
    f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=8, rms=1e-3, seed=1)
    with BlaGPU.from_ufield(f, precision=SIMSON_MIXED if form == "mixed" else SIMSON_FP64) as g:
        g.upload_velocity(f.u, -2.0 / RE)
        g.run(1, 0.01, icfl=0)
        p1 = g.pressure()
        g.run(9, 0.01, icfl=0, first=False)
        p10 = g.pressure()
    for p, key in ((p1, "p1"), (p10, "p10")):
        e, odd = masked_rel(p, golden_p[key])
        print(f"    {key}: rel L2 {e:.2e}, oddball {odd:.2e}", flush=True)
        # mixed: FP32 state (velocities to ~1e-7) and FP32 products
        assert e < (1e-5 if form == "mixed" else 1e-9), (key, e)
        assert odd < (1e-5 if form == "mixed" else 1e-9), (key, odd)


@pytest.mark.gpu
def test_gpu_pressure_matches_oracle_64():
    from simson_amd.gpu import BlaGPU
    print("[gpu] pressure vs oracle 64x65x48", flush=True)
    sys.stdout.flush()
    nx, ny, nz = 64, 65, 48
    xl, zl = 4 * math.pi, 4 * math.pi / 3
    f = fields.channel_field(nx, ny, nz, xl, zl, 4200.0, kmax=10, rms=5e-2, seed=4)
    o = pyoracle.Oracle(nx, ny, nz, 4200.0, xl, zl)
    ur, ui, puw = o.new_state(f.u)
    o.run(ur, ui, puw, 2, 0.005, icfl=0)
    ref = o.pressure(ur, ui)
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / 4200.0)
        g.run(2, 0.005, icfl=0)
        p = g.pressure()
    d = np.delete(p - ref, nz // 2, axis=0)
    r = np.delete(ref, nz // 2, axis=0)
    e = float(np.linalg.norm(d) / np.linalg.norm(r))
    print(f"    rel L2 {e:.2e}", flush=True)
    assert e < 1e-9
