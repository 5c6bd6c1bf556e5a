"""CPU restatement of bla's per-plane amplitude sums (test infrastructure).

boxamp_ref follows bla/boxamp.f:52-200 (nfzsym = 0): for every y plane,
sums over the stored Fourier modes (kx = 0..nx/2-1, all kz except the
oddball -nz/2) with weight 2 for kx > 0 and 1 for kx = 0, leaving out the
(0,0) mode, plus the (0,0) values.  Only the tests and the benchmark's CPU
leg use this module, as the checker of the GPU statistics.

Pinned by tests/test_stats.py against the amp.stat table the reference bla
wrote for the golden 32x33x32 run (fields after 1 and 10 steps).
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
