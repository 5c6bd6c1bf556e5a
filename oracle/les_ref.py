"""ADM-RT filter operators of bla (iles = 1), derived from their defining
conditions.  Test infrastructure: the oracle's statement of the filters.

The filters are those of Stolz, Adams and Kleiser, "An approximate
deconvolution model for large-eddy simulation with application to
incompressible wall-bounded flows", Phys. Fluids 13 (2001) 997, as bla
uses them (init_filter.f:9-361 builds them, hpf_y.f applies the wall-normal
one, bla.f:1494-1518 forms the x-z transfer function).

Each row filter is a five-point stencil with explicit weights b_j and
implicit weights a_j at grid offsets j, G = B/A with
B(k) = sum_j b_j e^{ijk}, A(k) = sum_j a_j e^{ijk}:

* explicit weights, at the signed distances p_j of the stencil points:
    sum b_j = 1, sum b_j p_j = 0, sum b_j p_j^2 = 0  (second-order moments),
    sum (-1)^j b_j = 0                                (B(pi) = 0),
  plus one condition for the fifth degree of freedom:
    interior (centred) rows: minimum dispersion, i.e. the least
      integral over [0, pi] of (Im B(k))^2 = (pi/2) sum_m (b_m - b_-m)^2;
    the rows next to the walls (one point on one side, three on the
      other): a zero third moment, sum b_j p_j^3 = 0;
* implicit weights a_j = r b_j (j != 0), a_0 = 1 - sum_{j!=0} a_j, so that
  A = 1 - r + r B; r is the smaller root of |B(kc)|^2 = |A(kc)|^2 / 4, the
  filter's half-transfer condition at kc = pi cutoff.

Here the explicit weights come from an SVD null-space parametrisation and a
one-dimensional minimisation (numpy); the library solves the same
conditions by another route, and both are pinned by the reference run
with sgs.i (tests/golden/les_32x33x32.npz).
"""
from __future__ import annotations

import numpy as np


def _constraints(p, j, third_moment):
    rows = [np.ones(5), p, p ** 2, (-1.0) ** j]
    rhs = [1.0, 0.0, 0.0, 0.0]
    if third_moment:
        rows.append(p ** 3)
        rhs.append(0.0)
    return np.array(rows), np.array(rhs)


def explicit_weights(p, j, third_moment):
    """b_j at signed distances p (5,) and integer offsets j (5,)."""
    p = np.asarray(p, float)
    j = np.asarray(j, int)
    Cm, d = _constraints(p, j, third_moment)
    if third_moment:
        return np.linalg.solve(Cm, d)
    # particular solution + t * null vector; minimise the odd part
    b0 = np.linalg.lstsq(Cm, d, rcond=None)[0]
    _, _, vt = np.linalg.svd(Cm)
    nv = vt[-1]
    # D b = (b_m - b_-m), m = 1, 2  (j is -2..2 here)
    D = np.zeros((2, 5))
    for m in (1, 2):
        D[m - 1, list(j).index(m)] = 1.0
        D[m - 1, list(j).index(-m)] = -1.0
    dn, db = D @ nv, D @ b0
    t = -float(dn @ db) / float(dn @ dn)
    return b0 + t * nv


def transfer(b, j, k):
    return np.exp(1j * np.outer(np.atleast_1d(k), j)) @ b


def implicit_factor(b, j, cutoff):
    B = transfer(b, j, np.pi * cutoff)[0]
    # |1 - r + r B|^2 = 4 |B|^2  ->  qa r^2 + qb r + qc = 0
    qa = abs(B - 1.0) ** 2
    qb = 2.0 * (B.real - 1.0)
    qc = 1.0 - 4.0 * abs(B) ** 2
    disc = qb * qb - 4.0 * qa * qc
    return float((-qb - np.sqrt(disc)) / (2.0 * qa))


def stencil(p, j, third_moment, cutoff):
    """(b, a): explicit and implicit weights of one row."""
    b = explicit_weights(p, j, third_moment)
    r = implicit_factor(b, j, cutoff)
    a = r * b
    c = list(j).index(0)
    a[c] = 1.0 - (a.sum() - a[c])
    return b, a


def homogeneous_transfer(m, cutoff):
    """G at k = pi t / (m - 1), t = 0 .. m-1, of the uniform stencil."""
    j = np.arange(-2, 3)
    b, a = stencil(j.astype(float), j, False, cutoff)
    k = np.pi * np.arange(m) / (m - 1)
    return np.real(transfer(b, j, k) / transfer(a, j, k))


def filter_operators(nx, ny, nz, cutoff, iord):
    """(lpf [nz][nx/2], L [ny][ny]): the x-z transfer function
    1 - (1 - Gx Gz)^N and the wall-normal low-pass I - H_y^N, N = iord + 1,
    H_y = A^-1 (A - B) on the Gauss-Lobatto points with zero wall rows."""
    if not 0.0 < cutoff < 1.0:
        raise ValueError("cutoff must lie in (0, 1): at 1 the transfer function "
                         "is undefined at the grid Nyquist wavenumber")
    N = iord + 1
    nxh = nx // 2
    gx = homogeneous_transfer(nxh + 1, cutoff)
    gz = homogeneous_transfer(nz // 2 + 1, cutoff)
    kz = np.array([z if z <= nz // 2 else nz - z for z in range(nz)])
    lpf = 1.0 - (1.0 - np.outer(gz[kz], gx[:nxh])) ** N
    eta = np.cos(np.pi * np.arange(ny) / (ny - 1))
    A = np.zeros((ny, ny))
    B = np.zeros((ny, ny))
    A[0, 0] = B[0, 0] = A[-1, -1] = B[-1, -1] = 1.0
    for k in range(1, ny - 1):
        if k == 1:
            cols = np.arange(0, 5)
        elif k == ny - 2:
            cols = np.arange(ny - 5, ny)
        else:
            cols = np.arange(k - 2, k + 3)
        j = cols - k
        p = eta[k] - eta[cols]
        b, a = stencil(p, j, k in (1, ny - 2), cutoff)
        B[k, cols] = b
        A[k, cols] = a
    H = np.linalg.solve(A, A - B)
    L = np.eye(ny) - np.linalg.matrix_power(H, N)
    return lpf, L
