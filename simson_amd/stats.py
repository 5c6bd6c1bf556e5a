"""Host side of bla's statistics: amplitudes (amp.stat) and x-y statistics.

The per-y plane sums (boxamp.f) come from the GPU (simson_gpu_amp); this
module does what bla's host does with them: the Clenshaw-Curtis weights of
preprbl.f:78-87, the integration over y and the line of wampbl.f for the
channel flow types (fltype 1), wampbl.f:124-278.

The x-y statistics sums (boxxys.f, simson_gpu_xys) are written and read in
the unformatted file layout of wxys.f:83-135 (velocity statistics only:
scalar = 0, no correlations, no time series).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass

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


XYS_NAMES = ("u", "v", "w", "u2", "v2", "w2", "omx", "omy", "omz", "omx2", "omy2", "omz2",
             "uv", "uw", "vw")


@dataclass
class XysFile:
    re: float
    xl: float
    zl: float
    t: float
    nx: int
    ny: int
    nz: int
    fltype: int
    dstar: float
    sumw: float
    stats: np.ndarray      # [nxys][ny][nx], already divided by sumw


def _rec(f, payload: bytes) -> None:
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def _read_rec(f) -> bytes:
    (n,) = struct.unpack("<i", f.read(4))
    data = f.read(n)
    (n2,) = struct.unpack("<i", f.read(4))
    if n2 != n or len(data) != n:
        raise ValueError("corrupt Fortran record marker in xy-statistics file")
    return data


def write_xys(path: str, sums: np.ndarray, sumw: float, re: float, xl: float, zl: float,
              t: float, nz: int, fltype: int = 1, dstar: float = 1.0) -> None:
    """Write accumulated x-y statistics sums[nxys][ny][nx] divided by the
    sum of the weights, in the record layout of wxys.f:83-135."""
    nxys, ny, nx = sums.shape
    with open(path, "wb") as f:
        _rec(f, struct.pack("<di4d", re, 0, xl, zl, t, 0.0))
        _rec(f, b"A" + struct.pack("<4d", 0.0, 0.0, 0.0, 0.0))
        _rec(f, struct.pack("<4i", nx, ny, nz, 0))
        _rec(f, struct.pack("<id", fltype, dstar))
        _rec(f, struct.pack("<d3i", sumw, nxys, 0, 0))
        for i in range(nxys):
            _rec(f, (sums[i] * (1.0 / sumw)).astype("<f8").tobytes())


def read_xys(path: str) -> XysFile:
    """Read a velocity x-y statistics file written by wxys.f (fltype 1-5)."""
    with open(path, "rb") as f:
        re_, _pou, xl, zl, t, _ = struct.unpack("<di4d", _read_rec(f)[:44])
        _read_rec(f)                                   # 'A', mhd_n, b0
        nx, ny, nz, _nfzsym = struct.unpack("<4i", _read_rec(f))
        fltype, dstar = struct.unpack("<id", _read_rec(f)[:12])
        if fltype < 0 or fltype >= 6:
            _read_rec(f)
        sumw, nxys, _nxysth, _scalar = struct.unpack("<d3i", _read_rec(f)[:20])
        stats = np.empty((nxys, ny, nx))
        for i in range(nxys):
            stats[i] = np.frombuffer(_read_rec(f), dtype="<f8").reshape(ny, nx)
    return XysFile(re=re_, xl=xl, zl=zl, t=t, nx=nx, ny=ny, nz=nz, fltype=fltype,
                   dstar=dstar, sumw=sumw, stats=stats)
