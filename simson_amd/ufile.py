"""Simson ``.u`` velocity-field files (Fortran unformatted sequential).

Format (little-endian, 4-byte record markers), as written by
``bla/wdiscbl.f:50-136`` and read by ``bla/rdiscbl.f:93-411``:

* record 1: ``re`` (f8), ``pou`` (logical, 4 bytes, always .false.),
  ``xl``, ``zl``, ``t``, ``xs`` (f8)  -- 44 bytes without scalars
* record 2: ``nx``, ``nyp``, ``nzc``, ``nfzsym`` (i4)
* record 3: ``fltype`` (i4), ``dstar`` (f8)
* optional record 4 for boundary-layer flow types (``wdiscbl.f:59-65``)
* then for each component i (slowest), each z = 1..nzc, each y = 1..nyp
  (fastest): one record of ``nx`` f8 = nx/2 complex values (re, im
  interleaved) for kx = 0..nx/2-1.

Field values are Fourier coefficients in x and z at the physical
Gauss-Lobatto points y_j = cos(pi j/(nyp-1)), j = 0 being the upper wall.

In memory a field is a complex128 array ``u[comp, z, y, kx]`` with
``comp`` in (u, v, w); z index ``iz`` holds kz = iz for iz < nz/2 and
kz = iz - nz otherwise.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np


@dataclass
class UField:
    u: np.ndarray                 # complex128 [3, nz, nyp, nx//2]
    re: float
    xl: float
    zl: float
    t: float = 0.0
    xs: float = 0.0
    fltype: int = 1
    dstar: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def nx(self) -> int:
        return 2 * self.u.shape[3]

    @property
    def nyp(self) -> int:
        return self.u.shape[2]

    @property
    def nz(self) -> int:
        return self.u.shape[1]


def _rec(f, payload: bytes) -> None:
    n = len(payload)
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def _read_rec(f) -> bytes:
    hdr = f.read(4)
    if len(hdr) != 4:
        raise EOFError("truncated .u file")
    (n,) = struct.unpack("<i", hdr)
    data = f.read(n)
    (n2,) = struct.unpack("<i", f.read(4))
    if n2 != n or len(data) != n:
        raise ValueError("corrupt Fortran record marker in .u file")
    return data


def write_u(path: str, fld: UField) -> None:
    u = np.asarray(fld.u, dtype=np.complex128)
    ncomp, nz, nyp, nxh = u.shape
    with open(path, "wb") as f:
        _rec(f, struct.pack("<di4d", fld.re, 0, fld.xl, fld.zl, fld.t, 0.0))
        _rec(f, struct.pack("<4i", 2 * nxh, nyp, nz, 0))
        _rec(f, struct.pack("<id", fld.fltype, fld.dstar))
        if fld.fltype in (4, 5, 6):
            _rec(f, struct.pack("<2d", fld.extra.get("bstart", 0.0),
                                fld.extra.get("blength", 0.0)))
        elif 7 <= fld.fltype <= 9:
            _rec(f, struct.pack("<4d", fld.extra.get("bstart", 0.0),
                                fld.extra.get("blength", 0.0),
                                fld.extra.get("rlam", 0.0),
                                fld.extra.get("spanv", 0.0)))
        rows = u.view(np.float64).reshape(ncomp, nz, nyp, 2 * nxh)
        marker = struct.pack("<i", 16 * nxh)
        for i in range(ncomp):
            for z in range(nz):
                for y in range(nyp):
                    f.write(marker)
                    f.write(rows[i, z, y].astype("<f8").tobytes())
                    f.write(marker)


def read_u(path: str, ncomp: int = 3) -> UField:
    with open(path, "rb") as f:
        r1 = _read_rec(f)
        re_, _pou, xl, zl, t, xs = struct.unpack("<di4d", r1[:44])
        nx, nyp, nzc, nfzsym = struct.unpack("<4i", _read_rec(f))
        if nfzsym != 0:
            raise ValueError("symmetric (nfzsym=1) fields are not supported")
        r3 = _read_rec(f)
        fltype, dstar = struct.unpack("<id", r3[:12])
        extra = {}
        if fltype == -1:
            (extra["rlam"],) = struct.unpack("<d", _read_rec(f)[:8])
        elif fltype == -2:
            extra["rlam"], extra["spanv"] = struct.unpack("<2d", _read_rec(f)[:16])
        elif fltype in (4, 5, 6):
            extra["bstart"], extra["blength"] = struct.unpack("<2d", _read_rec(f)[:16])
        elif 7 <= fltype <= 9:
            vals = struct.unpack("<4d", _read_rec(f)[:32])
            extra.update(zip(("bstart", "blength", "rlam", "spanv"), vals))
        if fltype in (1, 2, 4, 5):
            dstar = 1.0
        u = np.empty((ncomp, nzc, nyp, nx // 2), dtype=np.complex128)
        for i in range(ncomp):
            for z in range(nzc):
                for y in range(nyp):
                    row = np.frombuffer(_read_rec(f), dtype="<f8")
                    u[i, z, y] = row[0::2] + 1j * row[1::2]
    return UField(u=u, re=re_, xl=xl, zl=zl, t=t, xs=xs, fltype=fltype,
                  dstar=dstar, extra=extra)


def to_split(u: np.ndarray, nslots: int) -> tuple[np.ndarray, np.ndarray]:
    """Place velocity components into main storage arrays ``ur, ui``.

    Main storage is ``ur(memnx, memny, memnz, memnxyz)`` in Fortran order
    (``bla/bla.f:67``), i.e. a C array ``[slot][z][y][x]``.
    """
    ncomp, nz, nyp, nxh = u.shape
    ur = np.zeros((nslots, nz, nyp, nxh), dtype=np.float64)
    ui = np.zeros_like(ur)
    ur[:ncomp] = u.real
    ui[:ncomp] = u.imag
    return ur, ui
