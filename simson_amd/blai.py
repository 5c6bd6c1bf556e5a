"""Reader and writer for ``bla.i`` runtime parameters (version 20070716).

The grammar follows ``bla/rparambl.f:160-1860``: one value per line, lines
starting with ``#`` are skipped (``bla/comment.f``), numbers and logicals are
read list-directed (first token of the line), file names with ``(a)``.
Conditional reads (``retau`` only when ``cflux`` is false, the amplitude
file name only when ``iamp > 0`` ...) are reproduced for the temporal
channel subset this library runs; any parameter that switches on a feature
outside that subset raises ``BlaInputError``.

The post-read adjustments of ``rparambl.f:1860-2033`` are applied: ``icfl``
rounded to multiples of ``nst``, ``maxit`` rounded down to a multiple of
``nst``, ``vart = dt <= 0`` and ``cflmax = cflmaxin * sqrt(3)`` for RK3.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

VERSION = 20070716


class BlaInputError(ValueError):
    """Raised for malformed or unsupported ``bla.i`` content.

    ``code`` is the reference stop code where one exists (13 for a bad
    parameter file, ``rparambl.f:2001``; 345431 for a wrong version).
    """

    def __init__(self, msg: str, code: int = 13):
        super().__init__(msg)
        self.code = code


@dataclass
class BlaParams:
    namnin: str = "init.u"
    namnut: str = "end.u"
    tmax: float = 0.0
    maxit: int = 0
    cpumax: float = 0.0
    wallmax: float = 0.0
    write_inter: bool = False
    dt: float = 0.01
    nst: int = 3
    cflmaxin: float = 0.8
    xl: float = 4 * math.pi
    varsiz: bool = False
    rot: float = 0.0
    cflux: bool = True
    retau: float = 0.0
    pert: bool = False
    ibc: int = 0
    cim: bool = False
    gall: bool = False
    suction: bool = False
    spat: bool = False
    cdev: float = 0.0
    # spatial simulation: fringe region (rparambl.f:547-720, ffun.f)
    fmax: float = 0.0
    fstart: float = 0.0
    fend: float = 0.0
    frise: float = 0.0
    ffall: float = 0.0
    sgs: bool = False
    isfd: int = 0
    imhd: int = 0
    loctyp: int = 0
    # loctyp = 3: F_i = amp_i exp(-(y/yscale)^2) f(x) sin(tomeg t)
    # (rparambl.f:908-935, locf.f:358-410)
    lamp: tuple = (0.0, 0.0, 0.0)
    lxscale: float = 1.0
    lxloc0: float = 0.0
    lyscale: float = 1.0
    lxtype: float = 0.0
    ltomeg: float = 0.0
    tripf: bool = False
    wbci: int = 0
    # wbci = 1: blowing and suction strip v(x, z, t) = amp f(x) cos(zbet z)
    # sin(tomeg t) at the lower wall, negative part scaled by damp
    # (rparambl.f:1206-1232, cwallbc.f:53-115)
    wamp: float = 0.0
    wdamp: float = 1.0
    wxstart: float = 0.0
    wxend: float = 0.0
    wxrise: float = 1.0
    wxfall: float = 1.0
    wzbet: float = 0.0
    wtomeg: float = 0.0
    icfl: int = 0
    iamp: int = 0
    namamp: str = "amp.stat"
    fileurms: bool = False
    longli: bool = False
    iext: int = 0
    ixys: int = 0
    namxys: str = "xy.stat"
    ixyss: int = 0
    txys: float = 0.0
    msave: int = 0
    tsave: list = field(default_factory=list)
    nmsave: list = field(default_factory=list)
    mwave: int = 0
    npl: int = 0

    # derived (rparambl.f:1957-2033)
    @property
    def vart(self) -> bool:
        return self.dt <= 0

    @property
    def cflmax(self) -> float:
        base = {1: 1.0, 3: math.sqrt(3.0), 4: math.sqrt(8.0)}[self.nst]
        return self.cflmaxin * base

    def adjusted(self) -> "BlaParams":
        p = BlaParams(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        if 0 < p.icfl < p.nst:
            p.icfl = p.nst
        if p.icfl > 0:
            p.icfl = min(p.icfl // p.nst * p.nst, 3 * p.nst)
        if p.dt <= 0 and p.icfl <= 0:
            p.icfl = p.nst
        p.maxit = max(p.maxit, 0) // p.nst * p.nst
        # rparambl.f:1496, 1511
        if p.ixys > 0:
            p.ixys = max(p.ixys, p.nst) // p.nst * p.nst
            p.ixyss = max(0, p.ixyss)
            if p.ixyss > 0:
                p.ixyss = max(p.ixyss, p.nst) // p.nst * p.nst
        return p


def _lines(text: str):
    for raw in text.splitlines():
        if raw.startswith("#"):
            continue
        yield raw


def _logical(tok: str) -> bool:
    t = tok.strip().lower().strip(".")
    if t.startswith("t"):
        return True
    if t.startswith("f"):
        return False
    raise BlaInputError(f"bad logical {tok!r}")


def _num(tok: str) -> float:
    return float(tok.lower().replace("d", "e"))


def parse(text: str) -> BlaParams:
    it = _lines(text)

    def nxt() -> str:
        for line in it:
            if line.strip() == "":
                # list-directed reads skip blank records
                continue
            return line
        raise BlaInputError("bla.i ended early")

    def tok() -> str:
        return nxt().split()[0]

    def name() -> str:
        return nxt().strip().split()[0]

    p = BlaParams()
    version = int(_num(tok()))
    if version != VERSION:
        raise BlaInputError(f"wrong bla.i version {version}", code=345431)
    p.namnin = name()
    p.namnut = name()
    p.tmax = max(_num(tok()), 0.0)
    p.maxit = max(int(_num(tok())), 0)
    p.cpumax = max(_num(tok()), 0.0)
    p.wallmax = max(_num(tok()), 0.0)
    p.write_inter = _logical(tok())
    p.dt = _num(tok())
    p.nst = int(_num(tok()))
    p.cflmaxin = _num(tok())
    p.xl = _num(tok())
    p.varsiz = _logical(tok())
    p.rot = _num(tok())
    p.cflux = _logical(tok())
    if not p.cflux:
        p.retau = _num(tok())
    p.pert = _logical(tok())
    if p.pert:
        raise BlaInputError("perturbation mode (pert) is not supported")
    p.ibc = int(_num(tok()))
    p.cim = _logical(tok())
    if p.cim:
        # setmat.f:32 aborts unconditionally in the reference as well
        raise BlaInputError("Chebyshev integration method is disabled", code=434342)
    p.gall = _logical(tok())
    p.suction = _logical(tok())
    if p.suction:
        raise BlaInputError("wall suction is not supported")
    p.spat = _logical(tok())
    if p.spat:
        # rparambl.f:547-720: only the fringe damping toward the computed
        # base flow; tabulated free stream, base flow files, forced waves,
        # OS modes, streaks and fringe waves are not supported
        if _logical(tok()):
            raise BlaInputError("tabulated free-stream velocity (tabfre) is not supported")
        if _logical(tok()):
            raise BlaInputError("base flow from a file (rbfl) is not supported")
        p.fmax, p.fstart, p.fend, p.frise, p.ffall = (_num(tok()) for _ in range(5))
        if _num(tok()) != 0.0 or _num(tok()) != 0.0:
            raise BlaInputError("forced waves in the fringe (ampob, amp2d) are not supported")
        if _logical(tok()) or _logical(tok()) or _logical(tok()):
            raise BlaInputError("OS modes, streaks and waves in the fringe are not supported")
        # rparambl.f:1898-1913
        if (p.fstart >= p.fend or p.frise <= 0 or p.ffall <= 0 or p.fmax <= 0 or
                abs(p.fstart) > .5 * p.xl or abs(p.fend) > .5 * p.xl):
            raise BlaInputError("fmax, frise, ffall must be positive and fstart < fend "
                                "within -xl/2 to xl/2")
    elif not p.suction:
        p.cdev = _num(tok())
    p.sgs = _logical(tok())          # LES parameters then come from sgs.i (simson_amd.sgsi)
    p.isfd = int(_num(tok()))
    if p.isfd != 0:
        raise BlaInputError("selective frequency damping is not supported")
    p.imhd = int(_num(tok()))
    if p.imhd != 0:
        raise BlaInputError("MHD is not supported")
    p.loctyp = int(_num(tok()))
    if p.loctyp == 3:
        vals = [_num(tok()) for _ in range(8)]
        p.lamp = tuple(vals[:3])
        p.lxscale, p.lxloc0, p.lyscale, p.lxtype, p.ltomeg = vals[3:]
        if p.lxtype not in (0.0, 1.0):
            raise BlaInputError("loctyp 3 needs x type 0 or 1")
    elif p.loctyp != 0:
        raise BlaInputError("only the localized forcing loctyp = 3 is supported")
    p.tripf = _logical(tok())
    if p.tripf:
        raise BlaInputError("trip forcing is not supported")
    p.wbci = int(_num(tok()))
    if p.wbci == 1:
        (p.wamp, p.wdamp, p.wxstart, p.wxend, p.wxrise, p.wxfall, p.wzbet,
         p.wtomeg) = (_num(tok()) for _ in range(8))
    elif p.wbci != 0:
        raise BlaInputError("only the blowing/suction strip wbci = 1 is supported")
    p.icfl = int(_num(tok()))
    p.iamp = int(_num(tok()))
    if p.iamp > 0:
        p.namamp = name()
        p.fileurms = _logical(tok())
    p.longli = _logical(tok())
    p.iext = int(_num(tok()))
    if p.iext > 0:
        name()
    p.ixys = max(0, int(_num(tok())))
    if p.ixys > 0:
        # rparambl.f:1504-1660 with mxys = 1 (bla.f:200): no phase average
        p.namxys = name()
        p.ixyss = int(_num(tok()))
        p.txys = _num(tok())
        if _logical(tok()) or _logical(tok()):
            raise BlaInputError("two-point correlations (corrf, corrf_x) are not supported")
        if _logical(tok()):
            raise BlaInputError("time series (serf) are not supported")
    p.msave = int(_num(tok()))
    if p.msave < 0 or p.msave == 100000:
        raise BlaInputError("periodic save modes are not supported")
    for _ in range(p.msave):
        p.tsave.append(_num(tok()))
        p.nmsave.append(name())
    p.mwave = max(0, int(_num(tok())))
    if p.mwave > 0:
        raise BlaInputError("wavenumber saving (mwave) is not supported")
    p.npl = max(0, int(_num(tok())))
    if p.npl > 0:
        raise BlaInputError("plane saving (npl) is not supported")
    if p.nst not in (3, 4):
        raise BlaInputError("nst must be 3 or 4")
    if p.xl <= 0:
        raise BlaInputError("xl must be > 0")
    if p.ibc not in (0, 100, 101):
        raise BlaInputError("only ibc=0 (no-slip walls), ibc=100 and ibc=101 (free stream: "
                            "the base flow's values or derivatives) are supported")
    if p.ibc in (100, 101) and p.cflux:
        raise BlaInputError(f"ibc={p.ibc} needs cflux = .false.")
    return p


def read(path: str) -> BlaParams:
    with open(path) as f:
        return parse(f.read())


def format_channel(p: BlaParams) -> str:
    """Write a channel ``bla.i`` that the reference ``bla`` accepts."""
    def L(v):
        return ".true." if v else ".false."
    out = [
        "# bla.i written by simson_amd",
        f"{VERSION}        Version of input file",
        p.namnin,
        p.namnut,
        f"{p.tmax!r}      Maximum simulation time (tmax)",
        f"{p.maxit}       Maximum iterations (maxit)",
        f"{p.cpumax!r}    Maximum CPU time (cpumax)",
        f"{p.wallmax!r}   Maximum wall-clock time (wallmax)",
        f"{L(p.write_inter)}  Save intermediate files (write_inter)",
        f"{p.dt!r}        Time step (dt)",
        f"{p.nst}         number of RK stages (nst)",
        f"{p.cflmaxin!r}  CFL number (cflmaxin)",
        f"{p.xl!r}        Box length (xl)",
        f"{L(p.varsiz)}   Variable size (varsiz)",
        f"{p.rot!r}       Rotation rate (rot)",
        f"{L(p.cflux)}    Constant mass flux (cflux)",
    ]
    if not p.cflux:
        out.append(f"{p.retau!r}   Re_tau")
    out += [
        ".false.         Perturbation mode (pert)",
        f"{p.ibc}         Boundary condition number (ibc)",
        ".false.         Chebyshev integration method (cim)",
        f"{L(p.gall)}    Galilei transformation (gall)",
        ".false.         Constant wall-suction (suction)",
        f"{L(p.spat)}    Spatial simulation (spat)",
    ]
    if p.spat:
        out += [".false.         Tabulated free-stream velocity (tabfre)",
                ".false.         Read base flow from file (rbfl)",
                f"{p.fmax!r}      Fringe strength (fmax)",
                f"{p.fstart!r}    Fringe start (fstart)",
                f"{p.fend!r}      Fringe end (fend)",
                f"{p.frise!r}     Fringe rise distance (frise)",
                f"{p.ffall!r}     Fringe fall distance (ffall)",
                "0.0             Amplitude oblique waves (ampob)",
                "0.0             Amplitude 2D TS waves (amp2d)",
                ".false.         Forcing with Orr-Sommerfeld modes (osmod)",
                ".false.         Forcing with streaks (streak)",
                ".false.         Forcing with waves (waves)"]
    else:
        out += [f"{p.cdev!r}      Temporal development vel. (cdev)"]
    out += [
        f"{L(p.sgs)}     Read sgs.i (sgs)",
        "0               SFD (isfd)",
        "0               MHD (imhd)",
        f"{p.loctyp}      Type of localized disturbance (loctyp)",
    ]
    if p.loctyp == 3:
        out += [f"{p.lamp[0]!r}  Amplitude x", f"{p.lamp[1]!r}  Amplitude y",
                f"{p.lamp[2]!r}  Amplitude z", f"{p.lxscale!r}  x scale",
                f"{p.lxloc0!r}  x position", f"{p.lyscale!r}  y scale",
                f"{p.lxtype!r}  x type", f"{p.ltomeg!r}  Frequency"]
    out += [
        ".false.         Trip forcing (tripf)",
        f"{p.wbci}       Boundary condition at lower wall (wbci)",
    ]
    if p.wbci == 1:
        out += [f"{p.wamp!r}  amplitude (wp1)", f"{p.wdamp!r}  damping of suction (wp2)",
                f"{p.wxstart!r}  xstart", f"{p.wxend!r}  xend", f"{p.wxrise!r}  xrise",
                f"{p.wxfall!r}  xfall", f"{p.wzbet!r}  spanwise wavenumber (zbet)",
                f"{p.wtomeg!r}  angular frequency (tomeg)"]
    out += [
        f"{p.icfl}        CFL calculation interval (icfl)",
        f"{p.iamp}        Amplitude calculation interval (iamp)",
    ]
    if p.iamp > 0:
        out += [p.namamp, f"{L(p.fileurms)}  Write urms velocities (fileurms)"]
    out += [
        ".false.         y-dependent statistics (longli)",
        "0               Extremum calculation interval (iext)",
        f"{p.ixys}       xy-statistics calculation interval (ixys)",
    ]
    if p.ixys > 0:
        out += [p.namxys, f"{p.ixyss}  xy-statistics saving interval (ixyss)",
                f"{p.txys!r}  xy-statistics start time (txys)",
                ".false.         Two-point correlations in z (corrf)",
                ".false.         Two-point correlations in x (corrf_x)",
                ".false.         Time series (serf)"]
    out += [
        f"{len(p.tsave)}  Number of 3D fields to save (msave)",
    ]
    for t, n in zip(p.tsave, p.nmsave):
        out += [f"{t!r}", n]
    out += [
        "0               Number of wavenumbers to save (mwave)",
        "0               Number of planes to save (npl)",
        "",
    ]
    return "\n".join(out)
