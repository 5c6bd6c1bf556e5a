"""CPU restatement of bla's boundary-layer base flows, their free-stream
data and the temporal layer's forcing (test infrastructure).

Only the tests use this module, as the checker of the library's host-side
base flow (simson_base_flow).  It follows:

- cubip.f       the cubic through the four table entries around x (the
                interval holding x in the middle, the end four near the
                table ends), the last entry beyond the table;
- xlim.f        x above xlow, else xlow (1/2 + r^3 - r^4/2), r = x/xlow,
                and xlow/2 for x <= 0;
- rbla.f:119    x0 = dstar^2 re (m + 1)/(2 dstar2^2);
- cbflow.f:161-240  U(x, y) = u(x1) f'(yb s(x1)) S + u(x2) f'(yb s(x2))
                (1 - S), yb = 1 + eta, x1 = xlim(xb + x0, x0/2), x2 =
                xlim(xb + x0 + xl, x0/2), xb = x wrapped into [bstart,
                bstart + xl), S = step((xb - bstart)/blength), u(x) =
                (x/x0)^m, s(x) = sqrt((m + 1) u(x) re/(2x));
- cbflow.f:241-300  V = -int_{-1}^{y} dU/dx: x Fourier coefficients (the
                Nyquist mode dropped), Chebyshev coefficients, times -i
                alpha, the Chebyshev integral of icheb.f with V(-1) = 0,
                back to physical points;
- bflow.f:80-105    dU_i/dy at y = +1 with the weights of preprbl.f:88-121
                (sum_k k^2 a_k), then x Fourier coefficients / nxp;
- cbflow.f:58-84 and nonlinbl.f:771-795  the temporal boundary layer's
                parallel profile and the mean force that keeps it from
                diffusing.

Here in numpy's FFTs and a dense cosine transform, independent of the
library's code.  Pinned by tests/test_blasius.py: the similarity profile's
closed forms (f''(0), the displacement thickness), continuity of (U, V),
the Blasius wall-normal velocity where the blending is complete, and the
reference bla's spatial Blasius run (tests/golden/blasius_64x33x8.npz,
which also holds the du0upp and x0 that bla prints), and the reference's
temporal Blasius run (tests/golden/tblasius_32x33x16.npz).
"""
from __future__ import annotations

import math

import numpy as np

from fringe_ref import step


def cubip(x: float, ftab: np.ndarray, dx: float) -> float:
    n = len(ftab)
    if x > dx * (n - 1):
        return float(ftab[-1])
    m = min(max(int(x / dx), 1), n - 3)     # 1-based start of the stencil
    p = 2.0 * x / dx - (2 * m + 1)
    f = ftab[m - 1:m + 3]
    nodes = (-3.0, -1.0, 1.0, 3.0)
    out = 0.0
    for i, xi in enumerate(nodes):
        w = 1.0
        for j, xj in enumerate(nodes):
            if j != i:
                w *= (p - xj) / (xi - xj)
        out += w * f[i]
    return out


def xlim(x: float, xlow: float) -> float:
    if x >= xlow:
        return x
    if x <= 0.0:
        return 0.5 * xlow
    r = x / xlow
    return xlow * (0.5 + r ** 3 - 0.5 * r ** 4)


def cheb_matrix(ny: int) -> np.ndarray:
    """T[k, j] = T_k(eta_j) = cos(pi j k/(ny-1))."""
    j = np.arange(ny)
    return np.cos(np.pi * np.outer(j, j) / (ny - 1))


def cheb_coeffs(u: np.ndarray) -> np.ndarray:
    """Chebyshev coefficients along axis 0 of values at eta_j."""
    ny = u.shape[0]
    c = np.ones(ny)
    c[0] = c[-1] = 0.5
    T = cheb_matrix(ny)
    return (c[:, None] * T * c[None, :]) @ u * (2.0 / (ny - 1))


def base_flow(nx: int, ny: int, xl: float, re: float, dstar: float, tab: np.ndarray,
              dstar2: float, bstart: float, blength: float, rlam: float = 0.0,
              xsc: float = 0.0):
    """(base [3][ny][nxp], bj complex [3][nx/2], du0upp) of fltype 6/7
    (no spanwise base velocity)."""
    nxp = 3 * nx // 2
    ymax = tab[-1, 0]
    dyb = ymax / (len(tab) - 1)
    fp = tab[:, 2]
    x0 = dstar ** 2 * re * (rlam + 1.0) / (dstar2 ** 2 * 2.0)
    eta = np.cos(np.pi * np.arange(ny) / (ny - 1))
    U = np.zeros((ny, nxp))
    for j in range(nxp):
        xc = (j - nxp // 2) / nxp * xl + xsc
        xb = xc - (int((xc - bstart) / xl + 1.0) - 1.0) * xl
        S = step((xb - bstart) / blength)
        xs = [xlim(xb + x0, x0 / 2), xlim(xb + x0 + xl, x0 / 2)]
        us = [(x / x0) ** rlam for x in xs]
        ss = [math.sqrt((rlam + 1.0) * u * re / (2.0 * x)) for u, x in zip(us, xs)]
        for y in range(ny):
            yb = 1.0 + eta[y]
            U[y, j] = (us[0] * cubip(yb * ss[0], fp, dyb) * S +
                       us[1] * cubip(yb * ss[1], fp, dyb) * (1.0 - S))
    # continuity
    Uh = np.fft.fft(U, axis=1)[:, :nxp // 2] / nxp
    alfa = 2.0 * np.pi / xl * np.arange(nxp // 2)
    a = cheb_coeffs(Uh) * (-1j * alfa)[None, :]
    b = np.zeros_like(a)
    k = np.arange(ny)
    b[2:ny - 1] = (a[1:ny - 2] - a[3:ny]) * (0.5 / k[2:ny - 1])[:, None]
    b[ny - 1] = a[ny - 2] * (0.5 / (ny - 1))
    b[1] = a[0] - 0.5 * a[2]
    sgn = np.where(k[1:] % 2 == 1, 1.0, -1.0)
    b[0] = (sgn[:, None] * b[1:]).sum(axis=0)
    Vh = cheb_matrix(ny) @ b                    # physical y, kx < nxp/2
    full = np.zeros((ny, nxp), complex)
    full[:, :nxp // 2] = Vh
    full[:, nxp // 2 + 1:] = np.conj(Vh[:, 1:][:, ::-1])
    V = np.fft.ifft(full, axis=1).real * nxp
    base = np.stack([U, V, np.zeros_like(U)])
    # dU_i/dy at +1: sum_k k^2 a_k, then x Fourier coefficients / nxp
    d = np.einsum("k,ckj->cj", k.astype(float) ** 2, cheb_coeffs(
        base.transpose(1, 0, 2).reshape(ny, -1)).reshape(ny, 3, nxp).transpose(1, 0, 2))
    bj = np.fft.fft(d, axis=1)[:, :nx // 2] / nxp
    return base, bj, float(bj[0, 0].real)


def temporal_base_flow(nx: int, ny: int, re: float, dstar: float, tab: np.ndarray,
                       dstar2: float, rlam: float = 0.0, cdev: float = 0.0, t: float = 0.0):
    """(base [3][ny][nxp], bj [3][nx/2], du0upp, force [3][ny]) of a temporal
    boundary layer without spanwise velocity (fltype 3, -1): cbflow.f:58-84
    (the profile at x0 with eta = (1 + y) sqrt(re/(2 x0)) at every x),
    bflow.f:80-105 for its derivatives at y = +1, and nonlinbl.f:771-795
    for the forcing at xblc = x0 + cdev t (without the factor nxp nzp)."""
    nxp = 3 * nx // 2
    ymax = tab[-1, 0]
    dyb = ymax / (len(tab) - 1)
    x0 = dstar ** 2 * re * (rlam + 1.0) / (dstar2 ** 2 * 2.0)
    eta = np.cos(np.pi * np.arange(ny) / (ny - 1))
    s = math.sqrt(re / (2.0 * x0))
    U = np.array([cubip((1.0 + e) * s, tab[:, 2], dyb) for e in eta])
    base = np.zeros((3, ny, nxp))
    base[0] = U[:, None]
    k = np.arange(ny)
    a = cheb_coeffs(U[:, None])[:, 0]
    du = float(np.sum(k.astype(float) ** 2 * a))
    bj = np.zeros((3, nx // 2), complex)
    bj[0, 0] = du
    xblc = x0 + cdev * t
    xx = (xblc / x0) ** rlam
    h1u = cdev * (rlam / x0) * (xblc / x0) ** (rlam - 1.0)
    h3u = 0.5 * (rlam + 1.0) / xblc
    force = np.zeros((3, ny))
    for y, e in enumerate(eta):
        ybl = 1.0 + e
        eb = ybl * math.sqrt((rlam + 1.0) * re * xx / (2.0 * xblc))
        h2u = (cdev * 0.5 * (rlam - 1.0) * xx ** 1.5 * math.sqrt(0.5 * (rlam + 1.0) * re)
               * (1.0 / xblc) ** 1.5 * ybl)
        force[0, y] = (cubip(eb, tab[:, 2], dyb) * h1u + cubip(eb, tab[:, 3], dyb) * h2u
                       - cubip(eb, tab[:, 4], dyb) * xx * xx * h3u)
    return base, bj, du, force
