"""Boundary-layer base flows of spatial simulations (host side).

The similarity table and the base flow are computed by the library's host
code (simson_fsc_table, simson_base_flow); this module wraps them and reads
and writes the table file bla uses, fsc.dat (bls/fsc.f:93-118, rbla.f:
24-125): a Fortran sequential unformatted file with the header record
(0., ymax, n, m, dstar2, nscalar) -- reals of 8 bytes, integers of 4 --
and one record of (eta, f, f', f'', f''', g, g', g'') per row.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import struct

import numpy as np

from .gpu import SimsonBaseFlow, SimsonError, lib
from .ufile import _read_rec, _rec


@dataclasses.dataclass
class FscTable:
    """Rows [n][8] = eta, f, f', f'', f''', g, g', g''; m; dstar2 the
    displacement thickness in eta (ymax - f(ymax))."""
    tab: np.ndarray
    m: float
    dstar2: float

    @property
    def ymax(self) -> float:
        return float(self.tab[-1, 0])


def fsc_table(m: float = 0.0, n: int = 20001, ymax: float = 30.0,
              eps: float = 1e-15) -> FscTable:
    """Falkner-Skan-Cooke table of bls/fsc.f (defaults: its fsc.i values)."""
    tab = np.zeros((n, 8))
    d2 = C.c_double()
    st = lib().simson_fsc_table(m, n, ymax, eps, tab.reshape(-1), C.byref(d2))
    if st != 0:
        raise SimsonError(st, "similarity solution failed (check m, n, ymax)")
    return FscTable(tab, m, d2.value)


def write_fsc(path: str, t: FscTable) -> None:
    with open(path, "wb") as f:
        _rec(f, struct.pack("<ddiddi", 0.0, t.ymax, t.tab.shape[0], t.m, t.dstar2, 0))
        for row in t.tab:
            _rec(f, struct.pack("<8d", *row))


def read_fsc(path: str) -> FscTable:
    """fsc.dat without scalars (rbla.f rejects a table that does not start
    at eta = 0)."""
    with open(path, "rb") as f:
        a, b, n, m, d2, nsc = struct.unpack("<ddiddi", _read_rec(f)[:40])
        if nsc != 0:
            raise ValueError("fsc.dat with passive scalars is not supported")
        if a != 0.0:
            raise ValueError("fsc.dat must start at eta = 0")
        tab = np.array([struct.unpack("<8d", _read_rec(f)[:64]) for _ in range(n)])
    if abs(tab[-1, 0] - b) > 1e-12 * b:
        raise ValueError("fsc.dat rows do not end at ymax")
    return FscTable(tab, m, d2)


def spatial_base_flow(nx: int, ny: int, xl: float, re: float, dstar: float,
                      table: FscTable, bstart: float, blength: float, fltype: int = 6,
                      spanv: float = 0.0, xsc: float = 0.0):
    """The base flow [3][ny][3nx/2] of cbflow.f (fltype 6, 7, 8) in the
    units of the .u file, and the free-stream data of ibc = 101: bj complex
    [3][nx/2] (dU_i/dy at y = +1 per kx) and du0upp."""
    p = SimsonBaseFlow(nx, ny, fltype, xl, re, dstar, table.dstar2, bstart, blength, table.m,
                       spanv, xsc, 0.0, 0.0)
    base = np.zeros((3, ny, 3 * nx // 2))
    bjr = np.zeros((3, nx // 2))
    bji = np.zeros_like(bjr)
    du = C.c_double()
    tab = np.ascontiguousarray(table.tab, np.float64)
    st = lib().simson_base_flow(C.byref(p), tab.reshape(-1), tab.shape[0], base.reshape(-1),
                                bjr.ctypes.data_as(C.c_void_p), bji.ctypes.data_as(C.c_void_p),
                                C.byref(du))
    if st != 0:
        raise SimsonError(st, "invalid base-flow parameters")
    return base, bjr + 1j * bji, du.value


def temporal_base_flow(nx: int, ny: int, xl: float, re: float, dstar: float,
                       table: FscTable, fltype: int = 3, spanv: float = 0.0,
                       cdev: float = 0.0, t: float = 0.0):
    """Temporal boundary layer (fltype 3, -1, -2) in the units of the .u
    file: the parallel base flow [3][ny][3nx/2], its free-stream data (bj,
    du0upp) and the mean force [3][ny] of nonlinbl.f:771-795 at time t."""
    p = SimsonBaseFlow(nx, ny, fltype, xl, re, dstar, table.dstar2, 0.0, 0.0, table.m,
                       spanv, 0.0, 0.0, 0.0)
    base = np.zeros((3, ny, 3 * nx // 2))
    bjr = np.zeros((3, nx // 2))
    bji = np.zeros_like(bjr)
    du = C.c_double()
    tab = np.ascontiguousarray(table.tab, np.float64)
    st = lib().simson_base_flow(C.byref(p), tab.reshape(-1), tab.shape[0], base.reshape(-1),
                                bjr.ctypes.data_as(C.c_void_p), bji.ctypes.data_as(C.c_void_p),
                                C.byref(du))
    if st != 0:
        raise SimsonError(st, "invalid base-flow parameters")
    force = np.zeros((3, ny))
    st = lib().simson_parallel_force(C.byref(p), tab.reshape(-1), tab.shape[0], cdev, t,
                                     force.reshape(-1))
    if st != 0:
        raise SimsonError(st, "invalid forcing parameters")
    return base, bjr + 1j * bji, du.value, force


@dataclasses.dataclass
class SpatialCase:
    """A spatial boundary-layer run in the units of the .u file (lengths
    and times times dstar, re over dstar, rates over dstar): the start
    field, the base flow [3][ny][3nx/2] with its free-stream data, the
    fringe (fmax, fstart, fend, frise, ffall) and the lower-wall strip
    (amp, damp, xstart, xend, xrise, xfall, zbet, tomeg) or None, and dt."""
    field: object
    base: np.ndarray
    bj: np.ndarray
    du0upp: float
    fringe: tuple
    wall: tuple | None
    dt: float
    dstar: float


def tollmien_schlichting_case(nx: int = 512, ny: int = 97, nz: int = 64,
                              re: float = 400.0, xl: float = 1000.0, yl: float = 30.0,
                              zl: float = 50.0, fringe=(1.0, -100.0, 0.0, 60.0, 20.0),
                              bstart: float = -100.0, blength: float = 100.0,
                              strip=(1e-4, 1.0, 40.0, 60.0, 5.0, 5.0, 0.0),
                              freq: float = 100.0, noise: float = 1e-6, kmax: int = 8,
                              seed: int = 3, dt: float = 0.02) -> SpatialCase:
    """Spatial Blasius boundary layer forced by a blowing and suction strip
    at the reduced frequency F = omega nu/U^2 10^6 (a Tollmien-Schlichting
    wave), with seeded noise.  Arguments in displacement thicknesses of the
    inflow, where Re_delta* = re: the flow enters at x = 0 after the fringe
    (which, with the base-flow blending zone, sits at the end of the
    periodic box), strip = (amp, damp, xstart, xend, xrise, xfall, zbet)."""
    from .fields import boundary_layer_field
    ds = 2.0 / yl
    tab = fsc_table()
    base, bj, du = spatial_base_flow(nx, ny, xl * ds, re / ds, ds, tab, bstart * ds,
                                     blength * ds)
    f = boundary_layer_field(nx, ny, nz, xl * ds, zl * ds, re / ds, ds, base, bstart * ds,
                             blength * ds, kmax=kmax, rms=noise, seed=seed)
    fmax, fstart, fend, frise, ffall = fringe
    fr = (fmax / ds, fstart * ds, fend * ds, frise * ds, ffall * ds)
    wall = None
    if strip is not None:
        amp, damp, x0, x1, xr, xf, zb = strip
        omega = freq * re * 1e-6                     # per delta*/U
        wall = (amp, damp, x0 * ds, x1 * ds, xr * ds, xf * ds, zb / ds, omega / ds)
    return SpatialCase(f, base, bj, du, fr, wall, dt * ds, ds)
