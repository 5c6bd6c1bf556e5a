"""Host side of bla's amplitude statistics (amp.stat).

The per-y plane sums (boxamp.f) come from the GPU (simson_gpu_amp); this
module does what bla's host does with them: the Clenshaw-Curtis weights of
preprbl.f:78-87, the integration over y and the line of wampbl.f for the
channel flow types (fltype 1), wampbl.f:124-278.
"""
from __future__ import annotations

import math

import numpy as np

AMP_COLUMNS = ("t", "urms", "vrms", "wrms", "omx_rms", "omy_rms", "omz_rms", "omy2_k2",
               "uv", "e0", "hplus", "px", "Re_b", "tau1", "tau2", "delta99", "delta_star",
               "theta", "w_b")


def wint(ny: int) -> np.ndarray:
    """Integration weights on the Gauss-Lobatto points (preprbl.f:78-87)."""
    N = ny - 1
    y = np.arange(ny)
    w = np.full(ny, -0.5)
    for i in range(1, (ny + 1) // 2 - 1):
        w += np.cos(math.pi * (2 * i * y) / N) / ((2 * i) ** 2 - 1)
    w += 0.5 * np.cos(math.pi * y) / (N * N - 1)
    w *= -4.0 / N
    w[0] *= 0.5
    w[-1] *= 0.5
    return w


def gridy(ny: int, dstar: float = 1.0) -> np.ndarray:
    """Wall distance of the grid points from the lower wall (preprbl.f:57)."""
    return (1.0 + np.cos(math.pi * np.arange(ny) / (ny - 1))) / dstar


def amp_row(amp: np.ndarray, t: float, re: float, px: float, dstar: float = 1.0) -> np.ndarray:
    """One amp.stat line (19 values) from the per-y sums amp[20][ny] of a
    channel (wampbl.f: the fltype 1 branch)."""
    a = np.array(amp, dtype=np.float64, copy=True)
    ny = a.shape[1]
    wt = wint(ny)
    a[7] = -a[7]                  # Reynolds stress sign
    a[12] = a[8]                  # mean streamwise velocity
    s = np.zeros(14)
    for i in range(10):
        s[i] = np.sum(a[i] * wt)
    s[12] = np.sum(a[12] * wt)
    # boundary-layer thicknesses from the mean profile amp(13)
    gy = gridy(ny, dstar)
    d1 = 0.0
    for y in range(ny):
        if d1 == 0.0 and a[12, y] <= 0.99:
            if a[12, y] == 0.99 or y == 0:
                d1 = gy[y]
            else:
                d1 = gy[y] + (0.99 - a[12, y]) * (gy[y - 1] - gy[y]) / (a[12, y - 1] - a[12, y])
    d2 = np.sum((1.0 - a[12]) * wt) / dstar
    d3 = np.sum(a[12] * (1.0 - a[12]) * wt) / dstar
    for j in range(7):
        s[j] = math.sqrt(s[j] * 0.5)
    wb = s[9] * 0.5
    e0 = math.sqrt((s[8] + s[9]) * 0.5)
    tau1, tau2 = a[11, 0], a[11, ny - 1]
    hplus = math.sqrt((abs(tau1) + abs(tau2)) * re / 2.0)
    return np.array([t / dstar, s[0], s[1], s[2], s[3] * dstar, s[4] * dstar, s[5] * dstar,
                     s[6], s[7], e0, hplus, px, s[12] * re / 2.0, tau1, tau2, d1, d2, d3, wb])


def fortran_e(v: float, d: int = 16, w: int = 25, ed: int = 3) -> str:
    """Fortran Ew.dEe edit descriptor: 0.ddd...E+eee, right-justified."""
    if v == 0.0 or not math.isfinite(v):
        body = f"0.{'0' * d}E+{0:0{ed}d}" if v == 0.0 else str(v)
        return body.rjust(w)
    s = f"{abs(v):.{d - 1}e}"             # one digit before the point, d digits
    digits, ex = s.split("e")
    mant = digits.replace(".", "")
    e = int(ex) + 1
    sign = "-" if v < 0 else ""
    es = "+" if e >= 0 else "-"
    return f"{sign}0.{mant}E{es}{abs(e):0{ed}d}".rjust(w)


def format_row(row) -> str:
    """amp.stat line format of wampbl.f:276 (19(e25.16e3,tr1))."""
    return "".join(fortran_e(float(v)) + " " for v in row)
