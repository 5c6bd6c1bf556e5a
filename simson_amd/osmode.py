"""Orr-Sommerfeld modes of plane Poiseuille flow as bla initial fields.

The linear-stability check of the channel step: a small two-dimensional
Tollmien-Schlichting wave v = eps Re[v_hat(y) exp(i alpha (x - c t))] on
U = 1 - y^2 must grow and travel with the complex phase speed c of the least
stable Orr-Sommerfeld eigenvalue.  For Re = 10^4 (centreline velocity,
half-height) and alpha = 1 that is c = 0.23752649 + 0.00373967 i
(Orszag, J. Fluid Mech. 50 (1971) 689-703).

The eigenproblem
    [U (D^2 - a^2) - U'' - (D^2 - a^2)^2 / (i a Re)] v = c (D^2 - a^2) v,
    v = Dv = 0 at y = +-1,
is discretised by Chebyshev collocation on the Gauss-Lobatto points
y_j = cos(pi j / (n - 1)) (the bla grid, y = +1 first), with the four
boundary conditions replacing the first two and last two rows.
"""
from __future__ import annotations

import math

import numpy as np

from .ufile import UField


def cheb_diff(n: int):
    """Gauss-Lobatto points y_j = cos(pi j/(n-1)) and the collocation
    derivative matrix (Trefethen, Spectral Methods in MATLAB, cheb.m)."""
    N = n - 1
    y = np.cos(np.pi * np.arange(n) / N)
    c = np.ones(n)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(n)
    X = np.tile(y, (n, 1)).T
    dX = X - X.T
    D = np.outer(c, 1.0 / c) / (dX + np.eye(n))
    D -= np.diag(D.sum(axis=1))
    return y, D


def os_eigen(re: float, alpha: float, n: int):
    """Least stable Orr-Sommerfeld eigenvalue c and eigenfunction v_hat on
    the n Gauss-Lobatto points (normalised to max |v_hat| = 1, real at its
    maximum)."""
    import scipy.linalg

    y, D = cheb_diff(n)
    D2 = D @ D
    I = np.eye(n)
    L = D2 - alpha ** 2 * I
    U = np.diag(1.0 - y ** 2)
    Upp = -2.0 * I
    A = U @ L - Upp - (L @ L) / (1j * alpha * re)
    B = L.astype(complex)
    # clamped walls: v(+-1) = 0, Dv(+-1) = 0
    A[0], B[0] = I[0], 0.0
    A[1], B[1] = D[0], 0.0
    A[-1], B[-1] = I[-1], 0.0
    A[-2], B[-2] = D[-1], 0.0
    w, V = scipy.linalg.eig(A, B)
    ok = np.isfinite(w) & (np.abs(w) < 10.0)
    k = np.argmax(np.where(ok, w.imag, -np.inf))
    v = V[:, k]
    v = v / v[np.argmax(np.abs(v))]
    return complex(w[k]), v


def os_field(nx: int, ny: int, nz: int, xl: float, zl: float, re: float, eps: float = 1e-5,
             kx: int = 1) -> tuple[UField, complex, np.ndarray]:
    """Poiseuille flow plus eps times the least stable 2-D mode at the
    streamwise wavenumber alpha = 2 pi kx / xl, as a bla field.  Returns the
    field, c and v_hat."""
    alpha = 2.0 * math.pi * kx / xl
    c, v = os_eigen(re, alpha, ny)
    _, D = cheb_diff(ny)
    uh = 1j * (D @ v) / alpha            # continuity: i alpha u + Dv = 0
    u = np.zeros((3, nz, ny, nx // 2), dtype=complex)
    y = np.cos(np.pi * np.arange(ny) / (ny - 1))
    u[0, 0, :, 0] = 1.0 - y ** 2
    # a real field eps Re[f exp(i alpha x)] has the coefficient eps f / 2
    u[0, 0, :, kx] = 0.5 * eps * uh
    u[1, 0, :, kx] = 0.5 * eps * v
    return UField(u=u, re=re, xl=xl, zl=zl), c, v


def mode_amplitude(u: np.ndarray, v_hat: np.ndarray, kx: int = 1) -> complex:
    """Projection of the wall-normal velocity of mode kx on v_hat."""
    v = u[1, 0, :, kx]
    return complex(np.vdot(v_hat, v) / np.vdot(v_hat, v_hat))
