"""Host-side mirror of bla's step structure over the C ABI (ctypes).

``BlaGPU`` owns one device state and exposes the calls bla.f makes on the hot
path: ``prhs`` (bla.f:1562), ``substep`` (one RK substep = nonlinbl over all
y planes then linearbl over all z planes, bla.f:2251-2475) and the fixed-dt
driver ``run`` (bla.f:1465-1477, 2384-2397).  The library must be built
(``python -m simson_amd.build``); there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .ufile import UField

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "_lib", "libsimson_gpu.so")

SIMSON_FP64 = 0
SIMSON_MIXED = 1
NSLOT = 7

# Every entry point declared in include/simson_gpu.h.
EXPORTS = (
    "simson_gpu_init", "simson_gpu_upload", "simson_gpu_prhs", "simson_gpu_nonlin",
    "simson_gpu_linear", "simson_gpu_substep",
    "simson_gpu_run", "simson_gpu_cfl", "simson_gpu_download", "simson_gpu_amp", "simson_gpu_amp_noalloc",
    "simson_gpu_xys",
    "simson_gpu_pressure",
    "simson_gpu_les", "simson_les_filter", "simson_les_filter_secondary",
    "simson_gpu_closure", "simson_gpu_sync",
    "simson_gpu_fringe", "simson_fringe_function", "simson_gpu_wall", "simson_gpu_set_time",
    "simson_gpu_free_stream", "simson_fsc_table", "simson_base_flow",
    "simson_parallel_force", "simson_gpu_mean_force", "simson_gpu_local_force",
    "simson_local_force_pattern",
    "simson_wall_velocity",
    "simson_mlp_create", "simson_mlp_nparams", "simson_mlp_params", "simson_mlp_set_data",
    "simson_mlp_train", "simson_mlp_grad", "simson_mlp_infer", "simson_mlp_last_error",
    "simson_mlp_destroy",
    "simson_gpu_timing_enable", "simson_gpu_timing_read", "simson_gpu_last_error",
    "simson_gpu_finalize",
)


class SimsonParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("precision", C.c_int), ("ibc", C.c_int), ("cflux", C.c_int),
        ("re", C.c_double), ("xl", C.c_double), ("zl", C.c_double),
        ("rot", C.c_double), ("retau", C.c_double),
    ]


class SimsonLes(C.Structure):
    _fields_ = [("iles", C.c_int), ("iord", C.c_int), ("cutoff", C.c_double),
                ("chi", C.c_double), ("cs", C.c_double), ("ineg", C.c_int)]


class SimsonFringe(C.Structure):
    _fields_ = [("fmax", C.c_double), ("fstart", C.c_double), ("fend", C.c_double),
                ("frise", C.c_double), ("ffall", C.c_double), ("xsc", C.c_double)]


class SimsonWall(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("amp", "damp", "xstart", "xend", "xrise", "xfall",
                                         "zbet", "tomeg", "xsc", "zsc")]


class SimsonBaseFlow(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("fltype", C.c_int)] + [
        (n, C.c_double) for n in ("xl", "re", "dstar", "dstar2", "bstart", "blength", "rlam",
                                  "spanv", "xsc", "u0low", "w0low")]


class SimsonLocf(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("ampx", "ampy", "ampz", "xscale", "xloc0", "yscale",
                                         "xtype", "tomeg", "xsc")]


class SimsonError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"simson_gpu error {code}: {msg}")
        self.code = code


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"HIP library missing: {LIBPATH} (python -m simson_amd.build)")
        L = C.CDLL(LIBPATH)
        vp = C.c_void_p
        L.simson_gpu_init.argtypes = [C.POINTER(SimsonParams), C.POINTER(vp)]
        L.simson_gpu_upload.argtypes = [vp, _dp, _dp, C.c_int, C.c_void_p, C.c_double]
        L.simson_gpu_prhs.argtypes = [vp, C.c_double]
        L.simson_gpu_nonlin.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.simson_gpu_linear.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(C.c_double)]
        L.simson_gpu_substep.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.simson_gpu_run.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_int,
                                     C.POINTER(C.c_double)]
        L.simson_gpu_download.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                          C.POINTER(C.c_double)]
        L.simson_gpu_cfl.argtypes = [vp, C.POINTER(C.c_double)]
        L.simson_gpu_amp.argtypes = [vp, _dp]
        L.simson_gpu_amp_noalloc.argtypes = [vp, _dp]
        L.simson_gpu_xys.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.simson_gpu_pressure.argtypes = [vp, _dp, _dp]
        L.simson_gpu_les.argtypes = [vp, C.POINTER(SimsonLes)]
        L.simson_les_filter_secondary.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int,
                                                  _dp]
        L.simson_les_filter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                        _dp, _dp]
        L.simson_gpu_closure.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
        L.simson_gpu_sync.argtypes = [vp]
        L.simson_gpu_fringe.argtypes = [vp, C.POINTER(SimsonFringe), C.c_void_p]
        L.simson_gpu_wall.argtypes = [vp, C.POINTER(SimsonWall)]
        L.simson_gpu_free_stream.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_double]
        L.simson_fsc_table.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, _dp,
                                       C.POINTER(C.c_double)]
        L.simson_base_flow.argtypes = [C.POINTER(SimsonBaseFlow), _dp, C.c_int, _dp,
                                       C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        L.simson_parallel_force.argtypes = [C.POINTER(SimsonBaseFlow), _dp, C.c_int, C.c_double,
                                            C.c_double, _dp]
        L.simson_gpu_mean_force.argtypes = [vp, C.c_void_p]
        L.simson_gpu_local_force.argtypes = [vp, C.POINTER(SimsonLocf)]
        L.simson_local_force_pattern.argtypes = [C.c_int, C.c_int, C.c_double,
                                                 C.POINTER(SimsonLocf), _dp, _dp]
        L.simson_gpu_set_time.argtypes = [vp, C.c_double]
        L.simson_wall_velocity.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                           C.POINTER(SimsonWall), C.c_double, _dp, _dp]
        L.simson_fringe_function.argtypes = [C.c_int, C.c_double, C.POINTER(SimsonFringe), _dp]
        L.simson_gpu_timing_enable.argtypes = [vp, C.c_int]
        L.simson_gpu_timing_read.argtypes = [vp, C.POINTER(C.c_char_p),
                                             C.POINTER(C.c_double), C.POINTER(C.c_long),
                                             C.c_int]
        L.simson_gpu_last_error.restype = C.c_char_p
        L.simson_gpu_last_error.argtypes = [vp]
        L.simson_gpu_finalize.argtypes = [vp]
        for name in EXPORTS:
            getattr(L, name).restype = getattr(L, name).restype or C.c_int
        L.simson_gpu_finalize.restype = None
        _lib = L
    return _lib


class BlaGPU:
    """Device-resident bla state for one channel configuration."""

    def __init__(self, nx: int, ny: int, nz: int, re: float, xl: float, zl: float,
                 rot: float = 0.0, cflux: bool = True, retau: float = 0.0,
                 precision: int = SIMSON_FP64, ibc: int = 0):
        """precision: SIMSON_FP64, or SIMSON_MIXED (FP32 FFTs, products and
        Chebyshev transforms; FP64 tau solves, mean flow and partial RHS).
        ibc: 0 (no-slip walls), 100 or 101 (free stream above with the base
        flow's values or derivatives, see free_stream)."""
        self.nx, self.ny, self.nz = nx, ny, nz
        self.precision = precision
        p = SimsonParams(nx, ny, nz, precision, ibc, int(cflux), re, xl, zl, rot, retau)
        h = C.c_void_p()
        self._check(lib().simson_gpu_init(C.byref(p), C.byref(h)))
        self.h = h
        self.rot = rot

    @classmethod
    def from_ufield(cls, f: UField, xl: float | None = None, **kw) -> "BlaGPU":
        return cls(f.nx, f.nyp, f.nz, f.re, f.xl if xl is None else xl, f.zl, **kw)

    def _check(self, st: int):
        if st != 0:
            msg = lib().simson_gpu_last_error(getattr(self, "h", None) or None)
            raise SimsonError(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            lib().simson_gpu_finalize(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- data movement -------------------------------------------------
    def upload_velocity(self, u: np.ndarray, px: float):
        """Slots 0-2 from complex velocities u[comp, z, y, kx]."""
        ur = np.ascontiguousarray(u.real, dtype=np.float64)
        ui = np.ascontiguousarray(u.imag, dtype=np.float64)
        self._check(lib().simson_gpu_upload(self.h, ur, ui, 3, None, px))

    def upload_state(self, ur: np.ndarray, ui: np.ndarray, puw: np.ndarray, px: float):
        ur = np.ascontiguousarray(ur, dtype=np.float64)
        ui = np.ascontiguousarray(ui, dtype=np.float64)
        puw = np.ascontiguousarray(puw, dtype=np.float64)
        self._check(lib().simson_gpu_upload(self.h, ur, ui, NSLOT,
                                            puw.ctypes.data_as(C.c_void_p), px))

    def download(self, nslots: int = NSLOT):
        shape = (nslots, self.nz, self.ny, self.nx // 2)
        ur = np.zeros(shape)
        ui = np.zeros(shape)
        puw = np.zeros((2, self.ny))
        px = C.c_double()
        self._check(lib().simson_gpu_download(self.h, ur.ctypes.data_as(C.c_void_p),
                                              ui.ctypes.data_as(C.c_void_p), nslots,
                                              puw.ctypes.data_as(C.c_void_p), C.byref(px)))
        return ur, ui, puw, px.value

    def velocities(self) -> np.ndarray:
        ur, ui, _, _ = self.download(3)
        return ur + 1j * ui

    # --- the hot path ----------------------------------------------------
    def prhs(self, an: float):
        self._check(lib().simson_gpu_prhs(self.h, an))

    def nonlin(self, do_cfl=False) -> float:
        """Step 2 only (nonlinbl over all y planes); returns the CFL partial."""
        cfl = C.c_double(0.0)
        self._check(lib().simson_gpu_nonlin(self.h, int(do_cfl), C.byref(cfl)))
        return cfl.value

    def linear(self, an, bn, anp1, bnp1) -> float:
        """Step 3 only (linearbl over all z planes); returns px afterwards."""
        px = C.c_double(0.0)
        self._check(lib().simson_gpu_linear(self.h, an, bn, anp1, bnp1, C.byref(px)))
        return px.value

    def substep(self, an, bn, anp1, bnp1, do_cfl=False):
        cfl = C.c_double(0.0)
        px = C.c_double(0.0)
        self._check(lib().simson_gpu_substep(self.h, an, bn, anp1, bnp1, int(do_cfl),
                                             C.byref(cfl), C.byref(px)))
        return cfl.value, px.value

    def run(self, nsteps: int, dt: float, icfl: int = 0, first: bool = True) -> float:
        """nsteps fixed-dt RK3 steps; returns the last CFL*dt (0 if never)."""
        cfl = C.c_double(0.0)
        self._check(lib().simson_gpu_run(self.h, nsteps, dt, icfl, int(first), C.byref(cfl)))
        return cfl.value

    def cfl(self) -> float:
        """The boxcfl maximum of the current field (getdt.f), without
        stepping it; the CFL number of a step dt is (cfl pi + 2|rot|) dt."""
        c = C.c_double(0.0)
        self._check(lib().simson_gpu_cfl(self.h, C.byref(c)))
        return c.value

    def sync(self):
        self._check(lib().simson_gpu_sync(self.h))

    # --- timing ------------------------------------------------------------
    def timing(self, on: bool):
        self._check(lib().simson_gpu_timing_enable(self.h, int(on)))

    def les(self, iles=1, cutoff: float = 2.0 / 3.0, iord: int = 5, chi: float = 0.2,
            cs: float = 0.0, ineg: int = 1):
        """Large-eddy model of sgs.i: iles = 1 ADM-RT (cutoff = 1/cutoff_inv),
        iles = 2 Smagorinsky (cs = 0: dynamic procedure, clipping ineg),
        iles = 3 Smagorinsky on the high-pass filtered velocity (cutoff, iord,
        cs, ineg), 0 none.  iles may also be a simson_amd.sgsi.SgsParams."""
        if not isinstance(iles, int):
            sp = iles
            iles, cutoff, iord, chi = sp.iles, sp.cutoff, sp.iord, sp.chi
            cs, ineg = sp.cs, sp.ineg
        p = SimsonLes(iles, iord, cutoff, chi, cs, ineg)
        self._check(lib().simson_gpu_les(self.h, C.byref(p)))

    def fringe(self, fmax=None, fstart=0.0, fend=0.0, frise=0.0, ffall=0.0, base=None,
               xsc=0.0):
        """Fringe forcing toward the base flow base [3][ny][3nx/2] (physical
        points of the 3/2 x grid); fmax None switches it off."""
        if fmax is None:
            self._check(lib().simson_gpu_fringe(self.h, None, None))
            return
        f = SimsonFringe(fmax, fstart, fend, frise, ffall, xsc)
        b = np.ascontiguousarray(base, np.float64)
        self._check(lib().simson_gpu_fringe(self.h, C.byref(f), b.ctypes.data_as(C.c_void_p)))

    def wall(self, amp=None, damp=1.0, xstart=0.0, xend=0.0, xrise=1.0, xfall=1.0, zbet=0.0,
             tomeg=0.0, xsc=0.0, zsc=0.0):
        """Blowing and suction at the lower wall (bla.i wbci = 1); amp None
        switches it off."""
        if amp is None:
            self._check(lib().simson_gpu_wall(self.h, None))
            return
        w = SimsonWall(amp, damp, xstart, xend, xrise, xfall, zbet, tomeg, xsc, zsc)
        self._check(lib().simson_gpu_wall(self.h, C.byref(w)))

    def free_stream(self, bj=None, du0upp: float = 0.0):
        """ibc = 101 data: bj complex [3][nx/2], the base flow's dU_i/dy at
        y = +1 per kx (kz = 0 plane; None for zero), and du0upp; with
        ibc = 100 bj holds the base flow's values U_i at y = +1 instead."""
        if bj is None:
            self._check(lib().simson_gpu_free_stream(self.h, None, None, du0upp))
            return
        b = np.asarray(bj)
        br = np.ascontiguousarray(b.real, np.float64)
        bi = np.ascontiguousarray(b.imag, np.float64)
        self._check(lib().simson_gpu_free_stream(self.h, br.ctypes.data_as(C.c_void_p),
                                                 bi.ctypes.data_as(C.c_void_p), du0upp))

    def mean_force(self, force=None):
        """A mean force [3][ny] added to H's kx = kz = 0 mode every substep
        (temporal boundary layers); None switches it off."""
        if force is None:
            self._check(lib().simson_gpu_mean_force(self.h, None))
            return
        f = np.ascontiguousarray(force, np.float64)
        self._check(lib().simson_gpu_mean_force(self.h, f.ctypes.data_as(C.c_void_p)))

    def local_force(self, ampx=None, ampy=0.0, ampz=0.0, xscale=1.0, xloc0=0.0, yscale=1.0,
                    xtype=0.0, tomeg=0.0, xsc=0.0):
        """The localized volume force of bla.i loctyp = 3 (units of the .u
        file); ampx None switches it off."""
        if ampx is None:
            self._check(lib().simson_gpu_local_force(self.h, None))
            return
        p = SimsonLocf(ampx, ampy, ampz, xscale, xloc0, yscale, xtype, tomeg, xsc)
        self._check(lib().simson_gpu_local_force(self.h, C.byref(p)))

    def set_time(self, t: float):
        """Time of the start of the next substep (time-dependent forcing)."""
        self._check(lib().simson_gpu_set_time(self.h, t))

    def closure(self, params, in_scale=None, in_shift=None, yfeat=None, out_scale=None,
                out_shift=None):
        """Learned closure in the step (params None turns it off); see
        simson_amd.sgsnet.closure_maps for the feature maps."""
        if params is None:
            self._check(lib().simson_gpu_closure(self.h, None, None, None, None, None, None))
            return
        arrs = [np.ascontiguousarray(params, np.float32)] + [
            np.ascontiguousarray(a, np.float64)
            for a in (in_scale, in_shift, yfeat, out_scale, out_shift)]
        self._closure_keep = arrs
        self._check(lib().simson_gpu_closure(self.h, *[a.ctypes.data_as(C.c_void_p) for a in arrs]))

    def amp(self) -> np.ndarray:
        """Per-y plane statistics of boxamp.f, [20][ny] (row i = amp(:, i+1))."""
        a = np.zeros((20, self.ny))
        self._check(lib().simson_gpu_amp(self.h, a))
        return a

    def xys(self, dtn: float, out: np.ndarray | None = None, xs: float = 0.0,
            zs: float = 0.0) -> np.ndarray:
        """Add the x-y statistics 1-15 of boxxys.f for the current state,
        weighted by dtn, to out ([15][ny][nx], row i = xys(:, :, i+1));
        returns out (a new zero array when None)."""
        if out is None:
            out = np.zeros((15, self.ny, self.nx))
        if out.shape[1:] != (self.ny, self.nx) or out.shape[0] not in (12, 13, 14, 15):
            raise ValueError("out must be [nxys][ny][nx] with 12 <= nxys <= 15")
        if not (out.flags.c_contiguous and out.dtype == np.float64):
            raise ValueError("out must be a C-contiguous float64 array")
        self._check(lib().simson_gpu_xys(self.h, dtn, xs, zs, out.shape[0], out))
        return out

    def pressure(self) -> np.ndarray:
        """Pressure of the current state, complex [z][y][kx] at physical y
        (the slot bla writes to <name>.p)."""
        pr = np.zeros((self.nz, self.ny, self.nx // 2))
        pi = np.zeros_like(pr)
        self._check(lib().simson_gpu_pressure(self.h, pr, pi))
        return pr + 1j * pi

    def timing_read(self) -> dict:
        cap = 32
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_long * cap)()
        n = lib().simson_gpu_timing_read(self.h, names, ms, cnt, cap)
        if n < 0:
            self._check(n)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(n) if cnt[i] > 0}


def les_filter(nx: int, ny: int, nz: int, cutoff: float, iord: int):
    """The library's ADM-RT filter operators (lpf [nz][nx/2], L [ny][ny]);
    host computation, no device needed."""
    lpf = np.zeros((nz, nx // 2))
    L = np.zeros((ny, ny))
    st = lib().simson_les_filter(nx, ny, nz, cutoff, iord, lpf, L)
    if st != 0:
        raise SimsonError(st, "invalid ADM-RT filter parameters")
    return lpf, L


def les_filter_secondary(nx: int, nz: int, cutoff: float, iord: int):
    """The library's secondary x-z transfer function of the dynamic HPF model
    (lpf2 [nz][nx/2]); host computation, no device needed."""
    lpf2 = np.zeros((nz, nx // 2))
    st = lib().simson_les_filter_secondary(nx, nz, cutoff, iord, lpf2)
    if st != 0:
        raise SimsonError(st, "invalid LES filter parameters")
    return lpf2


def fringe_function(nx: int, xl: float, fmax: float, fstart: float, fend: float,
                    frise: float, ffall: float, xsc: float = 0.0) -> np.ndarray:
    """The library's fringe function lambda at the 3nx/2 physical points of
    the 3/2 x grid; host computation, no device needed."""
    lam = np.zeros(3 * nx // 2)
    f = SimsonFringe(fmax, fstart, fend, frise, ffall, xsc)
    st = lib().simson_fringe_function(nx, xl, C.byref(f), lam)
    if st != 0:
        raise SimsonError(st, "invalid fringe parameters")
    return lam


def wall_velocity(nx: int, nz: int, xl: float, zl: float, tc: float, amp: float,
                  damp: float = 1.0, xstart: float = 0.0, xend: float = 0.0, xrise: float = 1.0,
                  xfall: float = 1.0, zbet: float = 0.0, tomeg: float = 0.0) -> np.ndarray:
    """The library's spectral lower-wall v [nz][nx/2] at time tc (host)."""
    vr = np.zeros((nz, nx // 2))
    vi = np.zeros_like(vr)
    w = SimsonWall(amp, damp, xstart, xend, xrise, xfall, zbet, tomeg, 0.0, 0.0)
    st = lib().simson_wall_velocity(nx, nz, xl, zl, C.byref(w), tc, vr, vi)
    if st != 0:
        raise SimsonError(st, "invalid wall parameters")
    return vr + 1j * vi


def rk3_weights(dt: float):
    """Wray RK3 weights a, b, c (rkcoeff.f:21-32) scaled by dt."""
    a = (8. / 15., 5. / 12., 3. / 4.)
    b = (0., -17. / 60., -5. / 12.)
    return [dt * x for x in a], [dt * x for x in b]


def local_force_pattern(nx: int, ny: int, xl: float, ampx: float, ampy: float, ampz: float,
                        xscale: float, xloc0: float, yscale: float, xtype: float,
                        tomeg: float = 0.0, xsc: float = 0.0) -> np.ndarray:
    """The library's loctyp = 3 force pattern, complex [3][ny][nx/2] (host)."""
    pr = np.zeros((3, ny, nx // 2))
    pi = np.zeros_like(pr)
    p = SimsonLocf(ampx, ampy, ampz, xscale, xloc0, yscale, xtype, tomeg, xsc)
    st = lib().simson_local_force_pattern(nx, ny, xl, C.byref(p), pr, pi)
    if st != 0:
        raise SimsonError(st, "invalid localized-force parameters")
    return pr + 1j * pi
