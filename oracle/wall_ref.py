"""CPU restatement of bla's blowing and suction strip at the lower wall,
bla.i wbci = 1 (test infrastructure).

Only the tests and the benchmark's CPU leg use this module, as the checker
of the library's wall forcing.  It follows cwallbc.f:53-115 and 380-407:
on the physical points of the 3/2 grid (x_j = -xl/2 + j xl/nxp, z_k =
-zl/2 + k zl/nzp)

    v_w = max(s, s damp),  s = amp f(x) cos(zbet z) sin(tomeg tc)

(sin(tomeg tc) replaced by 1 when tomeg = 0), with f(x) = S((x - xstart)
/xrise) - S((x - xend)/xfall + 1) (S the smooth step of step.f; xend and
xstart taken into the box first), then the forward transforms vrfftf (nxp)
and vcfftf (nzp) divided by nxp nzp and truncated to the stored modes, the
z oddball zeroed.  linearbl.f:458-469 imposes v(-1) = v_w on the v solve of
every substep, with tc the time at the start of the substep (bla.f:1797).

Pinned by tests/test_wall.py against the reference bla's channel with the
strip (tests/golden/wall_32x33x32.npz).
"""
from __future__ import annotations

import numpy as np

from fringe_ref import step


def wall_v(nx: int, nz: int, xl: float, zl: float, amp: float, damp: float, xstart: float,
           xend: float, xrise: float, xfall: float, zbet: float, tomeg: float, tc: float,
           xsc: float = 0.0, zsc: float = 0.0) -> np.ndarray:
    """Spectral lower-wall v [nz][nx/2] (complex) at time tc."""
    nxp, nzp = 3 * nx // 2, 3 * nz // 2
    h = nxp // 2
    if xend >= xl / 2.0:
        xend -= xl
        if xstart >= xl / 2.0:
            xstart -= xl
    fx = np.zeros(nxp)
    for x in range(1, h + 1):
        xc11 = (2 * x - 1 - h - 1) / nxp * xl + xsc
        xc11 -= int((xc11 + xl / 2.0) / xl) * xl
        xc22 = (2 * x - h - 1) / nxp * xl + xsc
        xc22 -= int((xc22 + xl / 2.0) / xl) * xl
        fx[2 * (x - 1)] = step((xc11 - xstart) / xrise) - step((xc11 - xend) / xfall + 1.0)
        fx[2 * (x - 1) + 1] = step((xc22 - xstart) / xrise) - step((xc22 - xend) / xfall + 1.0)
    zc = zl * (np.arange(1, nzp + 1) - nzp // 2 - 1) / nzp + zsc
    st = 1.0 if tomeg == 0.0 else np.sin(tomeg * tc)
    s = amp * fx[None, :] * np.cos(zbet * zc)[:, None] * st      # [z][x]
    p = np.maximum(s, s * damp)
    X = np.fft.rfft(p, axis=1)[:, : nx // 2]
    X = np.fft.fft(X, axis=0) / (nxp * nzp)
    out = np.zeros((nz, nx // 2), dtype=complex)
    out[: nz // 2] = X[: nz // 2]
    out[nz // 2 + 1:] = X[nzp - nz + nz // 2 + 1:]
    return out


def run(o, ur, ui, puw, nsteps: int, dt: float, t0: float, wall) -> None:
    """nsteps RK3 steps of the oracle o with the lower-wall v_w = wall(tc)
    set before every linear step, tc = t + c_k dt at the start of substep k
    (bla.f:1797, 2392-2415); the weights as orc_run."""
    anrk = (8. / 15., 5. / 12., 3. / 4.)
    bnrk = (0., -17. / 60., -5. / 12.)
    cnrk = (0., 8. / 15., 2. / 3.)
    anp1, bnp1 = dt * anrk[0], 0.0
    o.prhs(ur, ui, puw, anp1)
    t = t0
    for it in range(1, 3 * nsteps + 1):
        tc = t + dt * cnrk[(it - 1) % 3]
        o.nonlin(ur, ui, do_cfl=False)
        o.set_wall(wall(tc))
        an, bn = anp1, bnp1
        anp1, bnp1 = dt * anrk[it % 3], dt * bnrk[it % 3]
        o.linear(ur, ui, puw, an, bn, anp1, bnp1)
        if it % 3 == 0:
            t += dt
    o.set_wall(None)
