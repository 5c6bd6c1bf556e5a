"""Training data for the learned subgrid closure: explicitly filtered DNS
fields of this solver, a-priori.

From a resolved channel field (main-storage layout: u, v, w as complex
[3][nz][ny][nx/2], Fourier in x and z, Gauss-Lobatto points in y) this
module forms, on the LES grid (every ``fw``-th point in x and z):

* the inputs: the nine components of the filtered velocity-gradient tensor
  g_ij = d ubar_i / dx_j and the wall distance y+ (in wall units, g+ = g nu /
  u_tau^2, y+ = (1 - |y|) Re_tau);
* the targets: the six independent subgrid stresses
  tau_ij = bar(u_i u_j) - ubar_i ubar_j (in wall units, tau / u_tau^2),
  ordered 11, 12, 13, 22, 23, 33 (the tau_ij of les.f / sgs_stress.f:62-131).

The filter is a top-hat of width fw grid spacings in x and z (transfer
function sin(k D/2) / (k D/2)), applied in Fourier space; the products
u_i u_j are formed on the 3/2-padded grid, so the filtered products carry
no aliasing.  x/z derivatives are spectral, y derivatives use the
Chebyshev differentiation matrix of the Gauss-Lobatto points.
"""
from __future__ import annotations

import numpy as np

from .fields import gl_points, kz_of_index

PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def cheb_diff(ny: int) -> np.ndarray:
    """Differentiation matrix D (du/dy = D u) on y_j = cos(pi j/(ny-1))."""
    n = ny - 1
    y = gl_points(ny)
    c = np.ones(ny)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(ny)
    dy = y[:, None] - y[None, :]
    D = np.outer(c, 1.0 / c) / (dy + np.eye(ny))
    D -= np.diag(D.sum(axis=1))
    return D


def _full_spectrum(u: np.ndarray, nx: int) -> np.ndarray:
    """[..., nz, ny, nx/2] storage (kx < nx/2) -> [..., ny, nz, nx/2 + 1]
    for numpy's inverse real transforms (the Nyquist kx column is zero)."""
    s = np.zeros(u.shape[:-3] + (u.shape[-2], u.shape[-3], nx // 2 + 1), complex)
    s[..., :nx // 2] = np.swapaxes(u, -3, -2)
    return s


def _synth(spec: np.ndarray, nxg: int, nzg: int, nz: int) -> np.ndarray:
    """Unnormalised synthesis u(x, z) = sum uhat exp(i(alpha x + beta z)) of
    [..., ny, nz, nx/2+1] spectra on an nzg x nxg grid (>= the storage
    size: zero padding)."""
    sh = spec.shape
    kz = kz_of_index(nz)
    pad = np.zeros(sh[:-2] + (nzg, nxg // 2 + 1), complex)
    idx = np.where(kz >= 0, kz, kz + nzg)
    pad[..., idx, :sh[-1]] = spec
    # numpy's inverse transforms carry 1/n: undo it
    return np.fft.irfft(np.fft.ifft(pad, axis=-2), n=nxg, axis=-1) * (nxg * nzg)


def _analyse(phys: np.ndarray, nx: int, nz: int) -> np.ndarray:
    """Inverse of _synth, truncated back to the storage modes."""
    nzg, nxg = phys.shape[-2:]
    f = np.fft.fft(np.fft.rfft(phys, axis=-1), axis=-2) / (nxg * nzg)
    kz = kz_of_index(nz)
    idx = np.where(kz >= 0, kz, kz + nzg)
    return f[..., idx, :nx // 2 + 1]


def tophat(nx: int, nz: int, xl: float, zl: float, fw: int) -> np.ndarray:
    """Transfer function [nz, nx/2+1] of the fw-point top-hat in x and z."""
    dx, dz = xl / nx, zl / nz
    ax = 2 * np.pi / xl * np.arange(nx // 2 + 1)
    bz = 2 * np.pi / zl * kz_of_index(nz)
    gx = np.sinc(ax * fw * dx / 2 / np.pi)
    gz = np.sinc(bz * fw * dz / 2 / np.pi)
    return gz[:, None] * gx[None, :]


def filtered(u: np.ndarray, nx: int, ny: int, nz: int, xl: float, zl: float, fw: int = 4):
    """Filtered velocity gradients g [9, ny, nz/fw, nx/fw] (order du/dx,
    du/dy, du/dz, dv/dx, ...) and subgrid stresses tau [6, ny, nz/fw, nx/fw]
    at the LES points (x = fw i dx, z = fw k dz), in the units of the field.
    The kz = -nz/2 row is not part of the velocity field (the solver's
    storage keeps the nonlinear term there, as bla does) and is dropped."""
    u = np.array(u, complex)
    u[:, nz // 2] = 0.0
    spec = _full_spectrum(u, nx)                     # [3, ny, nz, nx/2+1]
    G = tophat(nx, nz, xl, zl, fw)
    nxp, nzp = 3 * nx // 2, 3 * nz // 2
    up = _synth(spec, nxp, nzp, nz)                 # [3, ny, nzp, nxp]
    sl = (slice(None), slice(None), slice(None, None, fw), slice(None, None, fw))
    ub = spec * G                                   # filtered spectra
    ubar = _synth(ub, nx, nz, nz)[sl]               # [3, ny, nz/fw, nx/fw]
    tau = []
    for i, j in PAIRS:
        pij = _analyse(up[i] * up[j], nx, nz) * G
        tau.append(_synth(pij, nx, nz, nz)[sl[1:]] - ubar[i] * ubar[j])
    tau = np.stack(tau)
    # filtered velocity gradients
    ax = 2 * np.pi / xl * np.arange(nx // 2 + 1)
    bz = 2 * np.pi / zl * kz_of_index(nz)
    D = cheb_diff(ny)
    grads = []
    for i in range(3):
        grads.append(_synth(1j * ax[None, None, :] * ub[i], nx, nz, nz)[sl[1:]])
        grads.append(_synth(np.einsum("ab,bzx->azx", D, ub[i]), nx, nz, nz)[sl[1:]])
        grads.append(_synth(1j * bz[None, :, None] * ub[i], nx, nz, nz)[sl[1:]])
    return np.stack(grads), tau


def les_samples(u: np.ndarray, nx: int, ny: int, nz: int, xl: float, zl: float,
                re: float, fw: int = 4):
    """(X [n, 10], T [n, 6], re_tau) on the LES grid of one resolved field,
    in wall units of its mean wall shear."""
    g, tau = filtered(u, nx, ny, nz, xl, zl, fw)
    D = cheb_diff(ny)
    # wall units from the mean shear of the resolved field at both walls
    U = u[0, 0, :, 0].real
    dudy = D @ U
    tau_w = 0.5 * (abs(dudy[0]) + abs(dudy[-1])) / re
    utau = np.sqrt(tau_w)
    re_tau = utau * re
    y = gl_points(ny)
    yplus = (1.0 - np.abs(y)) * re_tau
    nzl, nxl = g.shape[-2:]
    X = np.concatenate([g.reshape(9, -1) / (utau * utau * re),
                        np.broadcast_to(yplus[:, None, None], (ny, nzl, nxl)).reshape(1, -1)])
    T = tau.reshape(6, -1) / (utau * utau)
    return X.T.copy(), T.T.copy(), re_tau


def smagorinsky_basis(X: np.ndarray, delta2_plus: float) -> np.ndarray:
    """The Smagorinsky form m_ij = -2 delta^2 |S| S_ij (wall units, the six
    components of les_samples) from the gradient inputs; the model's
    deviatoric stresses are c_s^2 m_ij."""
    g = X[:, :9].reshape(-1, 3, 3)
    S = 0.5 * (g + np.swapaxes(g, 1, 2))
    smag = np.sqrt(2.0 * np.einsum("nij,nij->n", S, S))
    t = -2.0 * delta2_plus * smag[:, None, None] * S
    return np.stack([t[:, i, j] for i, j in PAIRS], axis=1)
