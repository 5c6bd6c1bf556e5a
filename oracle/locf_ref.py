"""CPU restatement of bla's localized volume force loctyp = 3 (test
infrastructure).

Only the tests use this module, as the checker of the library's force.  It
follows locf.f:358-410: on the physical points of the 3/2 x grid (x_j =
-xl/2 + j xl/nxp + xsc, wrapped into the box) and at every z,

    F_i = amp_i exp(-((1 + y)/yscale)^2) f(x) sin(tomeg tc),
    f(x) = ((x - xloc0)/xscale)^xtype exp(-((x - xloc0)/xscale)^2),

added to the nonlinear term before its forward transforms (so, uniform in
z, to the kz = 0 plane: numpy's FFT over the nxp points divided by nxp,
truncated to kx < nx/2), with tc the time at the start of the substep
(bla.f:1797).  Pinned by tests/test_locf.py against the reference bla's
spatial Blasius run with this force (tests/golden/locf_64x33x8.npz).
"""
from __future__ import annotations

import numpy as np


def pattern(nx: int, ny: int, nz: int, xl: float, amp, xscale: float, xloc0: float,
            yscale: float, xtype: float, xsc: float = 0.0) -> np.ndarray:
    """The force before sin(tomeg tc): complex [3][nz][ny][nx/2], nonzero
    in the kz = 0 plane only."""
    nxp = 3 * nx // 2
    x = (np.arange(nxp) - nxp // 2) / nxp * xl + xsc
    x = x - np.trunc((x + xl / 2) / xl) * xl
    r = (x - xloc0) / xscale
    fx = (r ** xtype if xtype != 0 else np.ones_like(r)) * np.exp(-r * r)
    fk = np.fft.fft(fx)[:nx // 2] / nxp
    eta = np.cos(np.pi * np.arange(ny) / (ny - 1))
    fy = np.exp(-((1.0 + eta) / yscale) ** 2)
    out = np.zeros((3, nz, ny, nx // 2), complex)
    for c in range(3):
        out[c, 0] = amp[c] * fy[:, None] * fk[None, :]
    return out


def run(o, ur, ui, puw, nsteps: int, dt: float, t0: float, P: np.ndarray, tomeg: float,
        wall=None) -> None:
    """nsteps RK3 steps of the oracle o with the force sin(tomeg tc) P set
    before every nonlinear step (the oracle subtracts its external force),
    tc = t + c_k dt at the start of substep k; optionally the lower-wall
    v_w = wall(tc) as oracle/wall_ref.py sets it."""
    anrk = (8. / 15., 5. / 12., 3. / 4.)
    bnrk = (0., -17. / 60., -5. / 12.)
    cnrk = (0., 8. / 15., 2. / 3.)
    anp1, bnp1 = dt * anrk[0], 0.0
    o.prhs(ur, ui, puw, anp1)
    t = t0
    for it in range(1, 3 * nsteps + 1):
        tc = t + dt * cnrk[(it - 1) % 3]
        o.set_force(-np.sin(tomeg * tc) * P)
        o.nonlin(ur, ui, do_cfl=False)
        if wall is not None:
            o.set_wall(wall(tc))
        an, bn = anp1, bnp1
        anp1, bnp1 = dt * anrk[it % 3], dt * bnrk[it % 3]
        o.linear(ur, ui, puw, an, bn, anp1, bnp1)
        if it % 3 == 0:
            t += dt
    o.set_force(None)
    if wall is not None:
        o.set_wall(None)
