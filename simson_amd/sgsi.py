"""The LES parameter file ``sgs.i`` (read by rparamles.f:9-235 when bla.i
sets ``sgs``).

Grammar (version 20060909): comment lines start with ``#``; the values are
``version``, ``iles`` and, for ``iles = 1`` (ADM-RT, the relaxation-term
model), ``cutoff_inv``, ``iord`` and ``chi`` (rparamles.f:62-71); for
``iles = 2`` (Smagorinsky), ``cs`` and, when cs = 0 (the dynamic
procedure), ``ineg`` (rparamles.f:72-81); for ``iles = 3`` (the
Smagorinsky model on the high-pass filtered velocity), ``cutoff_inv``,
``iord``, ``cs`` and, when cs = 0, ``ineg`` (rparamles.f:82-91).  The defaults of
rparamles.f:28-51 apply to values a model does not read.  The checks of
rparamles.f:117-165 are repeated (cutoff = 1/cutoff_inv in [0, 1],
0 <= iord <= 5, 0 <= ineg <= 2, ineg = 1 for a fixed cs); like the
reference, a bad file stops with code 45543.

The step library supports every iles of the grammar (0 to 3).
"""
from __future__ import annotations

from dataclasses import dataclass

VERSION = 20060909


class SgsInputError(ValueError):
    def __init__(self, msg, code=45543):
        super().__init__(msg)
        self.code = code


@dataclass
class SgsParams:
    iles: int = 1
    cutoff_inv: float = 1.5
    iord: int = 5
    chi: float = 0.2
    cs: float = 0.0
    ineg: int = 1

    @property
    def cutoff(self) -> float:
        return 1.0 / self.cutoff_inv


def _values(text: str):
    for line in text.splitlines():
        t = line.strip()
        if not t or t.startswith("#"):
            continue
        yield t.split()[0]


def parse(text: str) -> SgsParams:
    it = _values(text)
    try:
        version = int(next(it))
        if version != VERSION:
            raise SgsInputError(f"wrong version of sgs.i: {version}", code=345432)
        p = SgsParams()
        p.iles = int(next(it))
        if p.iles == 1:
            p.cutoff_inv = float(next(it).replace("d", "e").replace("D", "e"))
            p.iord = int(next(it))
            p.chi = float(next(it).replace("d", "e").replace("D", "e"))
        elif p.iles == 2:
            p.cs = float(next(it).replace("d", "e").replace("D", "e"))
            if p.cs == 0.0:
                p.ineg = int(next(it))
        elif p.iles == 3:
            p.cutoff_inv = float(next(it).replace("d", "e").replace("D", "e"))
            p.iord = int(next(it))
            p.cs = float(next(it).replace("d", "e").replace("D", "e"))
            if p.cs == 0.0:
                p.ineg = int(next(it))
        elif p.iles != 0:
            raise SgsInputError(f"iles must be between 0 and 3: {p.iles}")
    except StopIteration:
        raise SgsInputError("sgs.i ends early") from None
    if not 0.0 <= p.cutoff <= 1.0:
        raise SgsInputError(f"cutoff wavenumber must be between 0 and 1: {p.cutoff}")
    if not 0 <= p.iord <= 5:
        raise SgsInputError(f"iord must be between 0 and 5: {p.iord}")
    if not 0 <= p.ineg <= 2:
        raise SgsInputError(f"ineg must be between 0 and 2: {p.ineg}")
    if p.cs != 0.0 and p.ineg != 1:
        raise SgsInputError(f"ineg must be 1: {p.ineg}")
    return p


def read(path: str) -> SgsParams:
    with open(path) as f:
        return parse(f.read())


def format_sgs(p: SgsParams) -> str:
    """Write an ``sgs.i`` that the reference ``bla`` accepts."""
    out = ["# sgs.i written by simson_amd", f"{VERSION}", f"{p.iles}"]
    if p.iles == 1:
        out += [f"{p.cutoff_inv!r}", f"{p.iord}", f"{p.chi!r}"]
    elif p.iles == 2:
        out += [f"{p.cs!r}"] + ([f"{p.ineg}"] if p.cs == 0.0 else [])
    elif p.iles == 3:
        out += [f"{p.cutoff_inv!r}", f"{p.iord}", f"{p.cs!r}"] + \
            ([f"{p.ineg}"] if p.cs == 0.0 else [])
    return "\n".join(out) + "\n"
