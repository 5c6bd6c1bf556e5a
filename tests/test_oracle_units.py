"""Pin each oracle building block against what the mathematics fixes.

Each check is chosen so that a plausible mistake (sign, index shift, a
dropped end-point factor, a wrong truncation row) fails it:
* DFTs against numpy.fft and the brute-force sums of vecfft.txt;
* Chebyshev transforms against their defining cosine sums and T_k(eta_j);
* dcheb/icheb against numpy's exact Chebyshev-series calculus;
* the tau solve against the defining tau equations (residual of
  u'' - lam u = f in T_0..T_{n-3} plus both wall values) and cosh solutions;
* whole steps against invariants: the Poiseuille fixed point, continuity,
  the constant mass flux.
"""
import math

import numpy as np
import numpy.polynomial.chebyshev as C
import pytest

import pyoracle
from simson_amd import fields

L = pyoracle.lib


@pytest.mark.parametrize("n", [2, 3, 6, 12, 48, 96, 192, 768, 45])
@pytest.mark.parametrize("sign", [-1, 1])
def test_cdft_matches_numpy(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    re, im = x.real.copy(), x.imag.copy()
    L().orc_cdft(n, re, im, sign)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert np.allclose(re + 1j * im, ref, atol=1e-13 * n, rtol=0)


@pytest.mark.parametrize("n", [6, 12, 48])
def test_real_dft_conventions(n):
    """vrfftf/vrfftb definitions, vecfft.txt:134-280 (brute force)."""
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n)
    xr = np.zeros(n // 2 + 1)
    xi = np.zeros(n // 2 + 1)
    L().orc_rdft_fwd(n, x, xr, xi)
    k = np.arange(n // 2 + 1)[:, None]
    m = np.arange(n)[None, :]
    assert np.allclose(xr, (x * np.cos(2 * np.pi * k * m / n)).sum(1), atol=1e-12)
    assert np.allclose(xi, -(x * np.sin(2 * np.pi * k * m / n)).sum(1), atol=1e-12)
    # synthesis: only Re X_0 and Re X_{n/2} enter
    Xr = rng.standard_normal(n // 2 + 1)
    Xi = rng.standard_normal(n // 2 + 1)
    out = np.zeros(n)
    L().orc_rdft_bwd(n, Xr, Xi, out)
    kk = np.arange(1, n // 2)[:, None]
    ref = (Xr[0] + Xr[n // 2] * (-1.0) ** np.arange(n)
           + 2 * (Xr[1:n // 2, None] * np.cos(2 * np.pi * kk * m / n)
                  - Xi[1:n // 2, None] * np.sin(2 * np.pi * kk * m / n)).sum(0))
    assert np.allclose(out, ref, atol=1e-12)


@pytest.mark.parametrize("n", [5, 9, 33, 97])
def test_chebyshev_transforms(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    f = np.zeros(n)
    L().orc_chbf(n, x, f)
    c = np.ones(n)
    c[0] = c[-1] = 0.5
    ang = np.pi * np.outer(np.arange(n), np.arange(n)) / (n - 1)
    assert np.allclose(f, c * (np.cos(ang) @ (c * x)), atol=1e-12)
    b = np.zeros(n)
    L().orc_chbb(n, f, b)
    assert np.allclose(b, x * (n - 1) / 2, atol=1e-11)      # vecfft.txt:518
    # chbb of a unit coefficient is T_k at the Gauss-Lobatto points
    e = np.zeros(n)
    e[3] = 1.0
    L().orc_chbb(n, e, b)
    eta = np.cos(np.pi * np.arange(n) / (n - 1))
    assert np.allclose(b, C.chebval(eta, e), atol=1e-13)


@pytest.mark.parametrize("n", [7, 33, 97])
def test_dcheb_icheb(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n)
    b = np.zeros(n)
    L().orc_dcheb(n, a, b)
    ref = C.chebder(a)
    assert np.allclose(b[: n - 1], ref, atol=1e-10 * n)
    assert b[n - 1] == 0.0
    # icheb integrates a series of degree n-2 exactly; b_0 is the constant
    a[n - 1] = 0.0
    I = np.zeros(n)
    L().orc_icheb(n, a, I, 0.25)
    assert I[0] == 0.25
    assert np.allclose(I[1:], C.chebint(a[: n - 1])[1:n], atol=1e-12)
    d = np.zeros(n)
    L().orc_dcheb(n, I, d)
    assert np.allclose(d[: n - 1], a[: n - 1], atol=1e-10)


@pytest.mark.parametrize("n", [9, 33, 97])
@pytest.mark.parametrize("lam", [0.0, 3.7, 1.1e6])
def test_tau_solve_satisfies_tau_equations(n, lam):
    rng = np.random.default_rng(n + int(lam))
    f = rng.standard_normal(n)
    bce, bco = 0.3, -0.7
    u = np.zeros(n)
    L().orc_tau_solve(n, lam, f, bce, bco, u)
    res = C.chebder(u, 2)[: n - 2] - lam * u[: n - 2] - f[: n - 2]
    scale = max(1.0, lam) * np.abs(u).max() + np.abs(f).max()
    assert np.abs(res).max() < 1e-12 * scale
    assert abs(C.chebval(1.0, u) - (bce + bco)) < 1e-12
    assert abs(C.chebval(-1.0, u) - (bce - bco)) < 1e-12


def test_tau_solve_cosh():
    """u'' - lam u = 0, u(+-1) = 1 has u = cosh(sqrt(lam) y)/cosh(sqrt(lam))."""
    n, lam = 65, 40.0
    u = np.zeros(n)
    L().orc_tau_solve(n, lam, np.zeros(n), 1.0, 0.0, u)
    y = np.linspace(-1, 1, 41)
    assert np.allclose(C.chebval(y, u), np.cosh(math.sqrt(lam) * y) / np.cosh(math.sqrt(lam)),
                       atol=1e-12)


def _run(nx, ny, nz, u0, re=3000.0, nsteps=5, dt=0.01, cflux=True):
    o = pyoracle.Oracle(nx, ny, nz, re, 4 * math.pi, 2 * math.pi, cflux=cflux)
    ur, ui, puw = o.new_state(u0)
    st, _ = o.run(ur, ui, puw, nsteps, dt, icfl=0)
    assert st == 0
    return o, ur, ui


def test_poiseuille_is_a_fixed_point():
    nx, ny, nz = 8, 17, 8
    f = fields.channel_field(nx, ny, nz, 4 * math.pi, 2 * math.pi, 3000.0,
                             kmax=0, rms=0.0, seed=1)
    o, ur, ui = _run(nx, ny, nz, f.u, nsteps=20)
    u = pyoracle.velocities(ur, ui)
    assert np.abs(u[0, 0, :, 0] - f.u[0, 0, :, 0]).max() < 1e-13
    assert np.abs(np.delete(u, nz // 2, axis=1)).sum() - np.abs(u[0, 0, :, 0]).sum() < 1e-12
    # px is re-estimated every substep from rounding-level flux deficits
    # (see the golden test); it stays at -2/Re to ~1e-9 relative
    assert abs(o.px - (-2.0 / 3000.0)) < 1e-8 * (2.0 / 3000.0)


def _cheb(ny, phys):
    """Chebyshev coefficients of a field given at the Gauss-Lobatto points."""
    out = np.zeros_like(phys)
    for idx in np.ndindex(phys.shape[:-1]):
        buf = np.zeros(ny)
        L().orc_chbf(ny, np.ascontiguousarray(phys[idx]), buf)
        out[idx] = buf * 2.0 / (ny - 1)
    return out


def test_continuity_and_mass_flux_after_steps():
    nx, ny, nz = 16, 25, 16
    f = fields.channel_field(nx, ny, nz, 4 * math.pi, 2 * math.pi, 3000.0,
                             kmax=5, rms=1e-2, seed=3)
    o, ur, ui = _run(nx, ny, nz, f.u, nsteps=10)
    u = pyoracle.velocities(ur, ui)
    alfa = 2 * np.pi / (4 * np.pi) * np.arange(nx // 2)
    kz = np.where(np.arange(nz) < nz // 2, np.arange(nz), np.arange(nz) - nz)
    beta = 2 * np.pi / (2 * np.pi) * kz
    chr_ = _cheb(ny, np.ascontiguousarray(u.real.transpose(0, 1, 3, 2)))
    chi = _cheb(ny, np.ascontiguousarray(u.imag.transpose(0, 1, 3, 2)))
    uc = chr_ + 1j * chi                          # [comp, z, x, k]
    dv = np.zeros_like(uc[1])
    for iz in range(nz):
        for ix in range(nx // 2):
            dv[iz, ix, : ny - 1] = C.chebder(uc[1, iz, ix])
    div = (1j * alfa[None, :, None] * uc[0] + dv + 1j * beta[:, None, None] * uc[2])
    div = np.delete(div, nz // 2, axis=0)
    assert np.abs(div).max() < 1e-11
    # constant flux: integral of the mean u over [-1,1] stays 4/3
    mean = uc[0, 0, 0].real
    flux = sum(-mean[k] * 2.0 / (k * k - 1) for k in range(0, ny, 2))
    assert abs(flux - 4.0 / 3.0) < 1e-12
