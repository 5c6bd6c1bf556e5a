"""CPU restatement of bla's statistics (test infrastructure).

boxamp_ref follows bla/boxamp.f:52-200 (nfzsym = 0): for every y plane,
sums over the stored Fourier modes (kx = 0..nx/2-1, all kz except the
oddball -nz/2) with weight 2 for kx > 0 and 1 for kx = 0, leaving out the
(0,0) mode, plus the (0,0) values.

boxxys_ref follows bla/boxxys.f:170-636 for the one-point statistics 1-15:
the stored planes with the z oddball zeroed (getxz.f:73-80, ipad = 0),
om_y = i beta u - i alpha w (boxxys.f:252-262), the lower-wall velocity
taken off the mean mode (:330-337), the frame shift of xzsh.f, the
unnormalised inverse z transform of length nz and real x transform of
length nx (vcfftb, vrfftb), and the z sums of u_i, u_i^2, om_i, om_i^2, uv,
uw, vw with weight dtn/nz (:540-636).

Only the tests and the benchmark's CPU leg use this module, as the checker
of the GPU statistics.  Pinned by tests/test_stats.py against the amp.stat
table and the xy-statistics file the reference bla wrote for the golden
32x33x32 runs.
"""
from __future__ import annotations

import numpy as np


def boxamp_ref(u: np.ndarray, omx: np.ndarray, omz: np.ndarray, alfa: np.ndarray,
               beta: np.ndarray) -> np.ndarray:
    """u[3, nz, ny, nxh], omx/omz[nz, ny, nxh] complex -> amp[20, ny]."""
    nz, ny, nxh = omx.shape
    omy = 1j * beta[:, None, None] * u[0] - 1j * alfa[None, None, :] * u[2]
    wt = np.ones(nxh)
    wt[1:] = 2.0
    mask = np.ones((nz, nxh))
    mask[nz // 2, :] = 0.0         # oddball
    mask[0, 0] = 0.0               # mean
    w = (mask * wt[None, :])[:, None, :]          # [nz, 1, nxh]
    k2 = alfa[None, :] ** 2 + beta[:, None] ** 2
    k2 = np.where(k2 == 0.0, 1.0, k2)[:, None, :]
    amp = np.zeros((20, ny))

    def s(f):
        return np.sum(w * f, axis=(0, 2))

    for i, f in enumerate((u[0], u[1], u[2], omx, omy, omz)):
        amp[i] = s(np.abs(f) ** 2)
    amp[6] = s(np.abs(omy) ** 2 / k2)
    amp[7] = s((u[0] * np.conj(u[1])).real)
    amp[8] = u[0, 0, :, 0].real
    amp[9] = u[2, 0, :, 0].real
    amp[10] = omx[0, :, 0].real
    amp[11] = omz[0, :, 0].real
    amp[13] = s((u[0] * np.conj(u[2])).real)
    amp[14] = s((u[1] * np.conj(u[2])).real)
    amp[15:18] = amp[0:3]
    return amp


def physical_xz(f: np.ndarray, nx: int) -> np.ndarray:
    """Stored plane spectra [nz, ny, nx/2] -> physical [nz, ny, nx] by the
    unnormalised backward transforms vcfftb (z, length nz) and vrfftb (x,
    length nx; only Re X_0 enters, no Nyquist term)."""
    nz = f.shape[0]
    g = np.fft.ifft(f, axis=0) * nz
    X = np.zeros(g.shape[:2] + (nx // 2 + 1,), dtype=complex)
    X[..., : g.shape[2]] = g
    return np.fft.irfft(X, n=nx, axis=2) * nx


def boxxys_ref(u: np.ndarray, omx: np.ndarray, omz: np.ndarray, alfa: np.ndarray,
               beta: np.ndarray, nx: int, dtn: float, u0low: float = 0.0,
               w0low: float = 0.0, xs: float = 0.0, zs: float = 0.0) -> np.ndarray:
    """u[3, nz, ny, nx/2], omx/omz[nz, ny, nx/2] complex -> xys[15, ny, nx]
    (row i = xys(:, :, i+1) of boxxys.f), one call's contribution."""
    nz = omx.shape[0]
    omy = 1j * beta[:, None, None] * u[0] - 1j * alfa[None, None, :] * u[2]
    flds = [u[0].copy(), u[1].copy(), u[2].copy(), omx.copy(), omy, omz.copy()]
    flds[0][0, :, 0] -= u0low
    flds[2][0, :, 0] -= w0low
    shift = np.exp(1j * (-xs * alfa[None, None, :] - zs * beta[:, None, None]))
    phys = []
    for f in flds:
        f = f * shift
        f[nz // 2] = 0.0
        phys.append(physical_xz(f, nx))
    c = dtn / nz
    out = np.zeros((15,) + phys[0].shape[1:])
    for i in range(3):
        out[i] = phys[i].sum(axis=0) * c
        out[3 + i] = (phys[i] ** 2).sum(axis=0) * c
        out[6 + i] = phys[3 + i].sum(axis=0) * c
        out[9 + i] = (phys[3 + i] ** 2).sum(axis=0) * c
    out[12] = (phys[0] * phys[1]).sum(axis=0) * c
    out[13] = (phys[0] * phys[2]).sum(axis=0) * c
    out[14] = (phys[1] * phys[2]).sum(axis=0) * c
    return out
