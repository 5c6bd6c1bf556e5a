"""CPU restatement of bla's fringe region for spatial simulations (test
infrastructure).

Only the tests and the benchmark's CPU leg use this module, as the checker
of the library's fringe.  It follows:

- step.f        the smooth step S(x) = 1/(1 + exp(1/(x-1) + 1/x)) on
                (0.02, 0.98], 0 below and 1 above;
- ffun.f:37-75  lambda at the physical points of the 3/2 x grid, x_j =
                -xl/2 + j xl/nxp + xsc (j = 2(i-1) is fring1(i), j + 1 is
                fring2(i)): fmax (S((x - fstart)/frise) - S((x - fend)/ffall
                + 1)) on the index range xstart..xend of the even points (the
                other points get no forcing), with the wrap-around case of a
                fringe across the box end;
- cbflow.f:117-131  the Poiseuille base flow of fltype 4, U = 1 - y^2.

Pinned by tests/test_fringe.py: the oracle stepped with this forcing
against the reference bla's spatial-channel run (tests/golden/
fringe_32x33x32.npz), and the step function's closed-form values.
"""
from __future__ import annotations

import math

import numpy as np

R = 0.499999999999999999          # ffun.f:24 (0.5 in double precision)


def step(x: float) -> float:
    if x <= 0.02:
        return 0.0
    if x <= 0.98:
        return 1.0 / (1.0 + math.exp(1.0 / (x - 1.0) + 1.0 / x))
    return 1.0


def fringe_lambda(nxp: int, xl: float, fmax: float, fstart: float, fend: float,
                  frise: float, ffall: float, xsc: float = 0.0) -> np.ndarray:
    """lambda [nxp] at the physical points of the 3/2 x grid."""
    lam = np.zeros(nxp)
    h = nxp // 2
    dnx = xl / nxp
    fstc = fstart + int((xsc - fstart) / xl + R) * xl
    fenc = fend + int((xsc - fend) / xl + R) * xl
    xstart = int(h * (fstc - xsc + xl / 2.0) / xl) + 1
    xend = int(h * (fenc - xsc + xl / 2.0) / xl) + 1
    wrap = xstart > xend
    imin, xendc = (0, h) if wrap else (1, 1)
    for i in range(imin, 2):
        xst = xstart ** i
        xen = max(xend, xendc ** i)
        for x in range(xst, xen + 1):          # 1-based even-point index
            xc1 = (2 * x - 1 - h - 1) * dnx + xsc
            xc2 = xc1 + dnx
            f1 = fmax * (step((xc1 - fstc) / frise) - step((xc1 - fenc) / ffall + 1.0))
            f2 = fmax * (step((xc2 - fstc) / frise) - step((xc2 - fenc) / ffall + 1.0))
            if wrap:
                f1 += fmax
                f2 += fmax
            lam[2 * (x - 1)] = f1
            lam[2 * (x - 1) + 1] = f2
    return lam


def poiseuille_base(ny: int, nxp: int) -> np.ndarray:
    """[3][ny][nxp]: u = 1 - eta^2, v = w = 0 (cbflow.f, fltype 1 and 4)."""
    eta = np.cos(math.pi * np.arange(ny) / (ny - 1))
    b = np.zeros((3, ny, nxp))
    b[0] = (1.0 - eta ** 2)[:, None]
    return b
