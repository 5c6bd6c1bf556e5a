"""ctypes wrapper of the CPU oracle (``oracle/_build/libblaref.so``).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the CPU
baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "_build", "libblaref.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"oracle not built: {LIBPATH} (run make -C oracle)")
        L = C.CDLL(LIBPATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, C.c_double]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_get_px.restype = C.c_double
        L.orc_get_px.argtypes = [C.c_void_p]
        L.orc_set_px.argtypes = [C.c_void_p, C.c_double]
        L.orc_prhs.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double]
        L.orc_nonlin.restype = C.c_double
        L.orc_nonlin.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.orc_linear.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double,
                                 C.c_double, C.c_double]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int,
                              C.POINTER(C.c_double)]
        L.orc_cdft.argtypes = [C.c_int, _dp, _dp, C.c_int]
        L.orc_rdft_bwd.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_rdft_fwd.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_chbf.argtypes = [C.c_int, _dp, _dp]
        L.orc_chbb.argtypes = [C.c_int, _dp, _dp]
        L.orc_dcheb.argtypes = [C.c_int, _dp, _dp]
        L.orc_icheb.argtypes = [C.c_int, _dp, _dp, C.c_double]
        L.orc_tau_solve.argtypes = [C.c_int, C.c_double, _dp, C.c_double, C.c_double, _dp]
        L.orc_tau_solve_neumann.argtypes = [C.c_int, C.c_double, _dp, C.c_double, C.c_double,
                                            _dp]
        L.orc_pressure.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_set_les.restype = C.c_int
        L.orc_set_les.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
        L.orc_set_force.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_fringe.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_wall.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_bc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double]
        L.orc_set_mean_force.argtypes = [C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


class Oracle:
    """One channel configuration; state arrays are owned by the caller."""

    NSLOT = 7

    def __init__(self, nx, ny, nz, re, xl, zl, rot=0.0, cflux=True, retau=0.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = lib().orc_create(nx, ny, nz, re, xl, zl, rot, int(cflux), retau)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def px(self):
        return lib().orc_get_px(self.h)

    @px.setter
    def px(self, v):
        lib().orc_set_px(self.h, v)

    def new_state(self, u: np.ndarray):
        """Main storage ur, ui ([slot][z][y][x]) and puw from velocities."""
        ur = np.zeros((self.NSLOT, self.nz, self.ny, self.nx // 2))
        ui = np.zeros_like(ur)
        ur[:3] = u.real
        ui[:3] = u.imag
        puw = np.zeros((2, self.ny))
        return ur, ui, puw

    def prhs(self, ur, ui, puw, an):
        lib().orc_prhs(self.h, ur, ui, puw, an)

    def nonlin(self, ur, ui, do_cfl=True):
        return lib().orc_nonlin(self.h, ur, ui, int(do_cfl))

    def linear(self, ur, ui, puw, an, bn, anp1, bnp1):
        lib().orc_linear(self.h, ur, ui, puw, an, bn, anp1, bnp1)

    def run(self, ur, ui, puw, nsteps, dt, icfl=0):
        cfl = C.c_double(0.0)
        st = lib().orc_run(self.h, ur, ui, puw, nsteps, dt, icfl, C.byref(cfl))
        return st, cfl.value

    def set_les(self, iles=1, cutoff=2.0 / 3.0, iord=5, chi=0.2):
        """ADM-RT relaxation model (sgs.i iles = 1) on, or off with iles = 0;
        the filter operators come from les_ref (their defining conditions)."""
        if iles == 0:
            self._les = None
            lib().orc_set_les(self.h, None, None, 0.0)
            return
        if iles != 1 or not 0 <= iord <= 5:
            raise ValueError("unsupported LES parameters")
        lpf, L = les_filter(self.nx, self.ny, self.nz, cutoff, iord)
        self._les = (np.ascontiguousarray(lpf), np.ascontiguousarray(L))
        lib().orc_set_les(self.h, self._les[0].ctypes.data, self._les[1].ctypes.data, chi)

    def set_fringe(self, lam=None, base=None):
        """Fringe forcing with lam [nxp] and base flow [3][ny][nxp]; None: off."""
        if lam is None:
            lib().orc_set_fringe(self.h, None, None)
            return
        self._fringe = (np.ascontiguousarray(lam, np.float64),
                        np.ascontiguousarray(base, np.float64))
        lib().orc_set_fringe(self.h, self._fringe[0].ctypes.data, self._fringe[1].ctypes.data)

    def set_bc(self, ibc=0, bj=None, du0upp=0.0):
        """Upper boundary: ibc 0 (wall) or 101 (free stream with the base
        flow's dU_i/dy at y = +1, bj complex [3][nx/2] for kz = 0, None for
        zero, and its mean du0upp)."""
        if bj is None:
            st = lib().orc_set_bc(self.h, ibc, None, None, du0upp)
        else:
            b = np.asarray(bj)
            self._bc = (np.ascontiguousarray(b.real, np.float64),
                        np.ascontiguousarray(b.imag, np.float64))
            st = lib().orc_set_bc(self.h, ibc, self._bc[0].ctypes.data, self._bc[1].ctypes.data,
                                  du0upp)
        if st != 0:
            raise ValueError(f"unsupported ibc {ibc}")

    def set_force(self, f=None):
        """External force, complex [3][nz][ny][nx/2] at physical y, which the
        nonlinear step subtracts (times nxp nzp) from H; None: off."""
        if f is None:
            lib().orc_set_force(self.h, None, None)
            return
        self._force = (np.ascontiguousarray(np.real(f), np.float64),
                       np.ascontiguousarray(np.imag(f), np.float64))
        lib().orc_set_force(self.h, self._force[0].ctypes.data, self._force[1].ctypes.data)

    def set_mean_force(self, f=None):
        """Mean force [3][ny] on the kx = kz = 0 mode of H; None: off."""
        if f is None:
            lib().orc_set_mean_force(self.h, None)
            return
        self._mf = np.ascontiguousarray(f, np.float64)
        lib().orc_set_mean_force(self.h, self._mf.ctypes.data)

    def set_wall(self, vw=None):
        """Lower-wall v_w [nz][nx/2] (complex spectral) for later linear
        steps; None: off."""
        if vw is None:
            lib().orc_set_wall(self.h, None, None)
            return
        self._wall = (np.ascontiguousarray(vw.real), np.ascontiguousarray(vw.imag))
        lib().orc_set_wall(self.h, self._wall[0].ctypes.data, self._wall[1].ctypes.data)

    def pressure(self, ur, ui):
        """Pressure [z][y][x] (complex) of the state in ur, ui (slots 0-4)."""
        pr = np.zeros((self.nz, self.ny, self.nx // 2))
        pi = np.zeros_like(pr)
        lib().orc_pressure(self.h, ur, ui, pr, pi)
        return pr + 1j * pi


def les_filter(nx, ny, nz, cutoff, iord):
    """(lpf [nz][nx/2], L [ny][ny]) of the ADM-RT filters (les_ref)."""
    import les_ref
    return les_ref.filter_operators(nx, ny, nz, cutoff, iord)


def velocities(ur, ui):
    return ur[:3] + 1j * ui[:3]
