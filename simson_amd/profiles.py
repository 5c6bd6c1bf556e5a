"""Channel statistics from Fourier-physical velocities, for comparing runs.

u has the layout of BlaGPU.velocities(): [3][nz][ny][nx/2] complex (the .u
file's Fourier-physical form, y at the Gauss-Lobatto points
y_j = cos(pi j/(ny-1)), kx >= 0).  The z-oddball plane kz = -nz/2 holds the
unnormalised nonlinear term in bla's storage, not velocity, and is left out.
Sums over kx > 0 count twice (the kx < 0 half is the conjugate).
"""
from __future__ import annotations

import numpy as np


def _drop_oddball(u):
    return np.delete(u, u.shape[1] // 2, axis=1)


def profiles(u):
    """Mean U, V, W (y) and the fluctuation statistics u_rms, v_rms, w_rms,
    uv (y) by Parseval.  Returns (mean[3][ny], rms[3][ny], uv[ny])."""
    u = _drop_oddball(u)
    mean = u[:, 0, :, 0].real.copy()
    w = np.ones(u.shape[3])
    w[1:] = 2.0
    p = np.abs(u) ** 2 * w[None, None, None, :]
    tot = p.sum(axis=(1, 3)) - np.abs(u[:, 0, :, 0]) ** 2
    rms = np.sqrt(np.maximum(tot, 0.0))
    cross = (u[0] * np.conj(u[1])).real * w[None, None, :]
    uv = cross.sum(axis=(0, 2)) - (u[0, 0, :, 0] * np.conj(u[1, 0, :, 0])).real
    return mean, rms, uv


def wall_slopes(U):
    """dU/dy at y = +1 and y = -1 of a profile on the Gauss-Lobatto points,
    from its Chebyshev interpolant: T_k'(1) = k^2, T_k'(-1) = (-1)^(k+1) k^2."""
    n = len(U)
    # interpolant coefficients by the discrete cosine sum on the
    # Gauss-Lobatto points (end points halved)
    j = np.arange(n)
    h = np.ones(n)
    h[0] = h[-1] = 0.5
    c = (2.0 / (n - 1)) * h * (np.cos(np.pi * np.outer(j, j) / (n - 1)) @ (h * np.asarray(U)))
    k2 = np.arange(n, dtype=float) ** 2
    sign = np.where(np.arange(n) % 2 == 0, -1.0, 1.0)
    return float((k2 * c).sum()), float((sign * k2 * c).sum())


def wall_shear(U, re):
    """Mean wall shear stress (|dU/dy| averaged over the two walls) / Re."""
    up, lo = wall_slopes(U)
    return 0.5 * (abs(up) + abs(lo)) / re


def deviation(a, b):
    """Largest difference over y, relative to the largest |b|."""
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def long_run(nx, ny, nz, re, xl, zl, dt, spinup, nsteps, every, seed=2, kmax=16, rms=0.1,
             controls=(1, 2), eps=1e-7, say=None):
    """The mixed form over a long run against FP64, with FP64 controls.

    A seeded field (rms over |kx|, |kz| <= kmax) is spun up for `spinup`
    FP64 steps.  From that state run `nsteps` steps each of the FP64 step,
    the mixed step and one FP64 control per seed in `controls`, whose
    velocities are perturbed by relative eps (Gaussian).  Statistics (mean
    U, u/v/w rms, uv, wall shear) are sampled every `every` steps and
    averaged over the second half of the run.  Returns a dict: for each
    statistic the deviation of the mixed average and the largest control
    deviation from the FP64 average, and the instantaneous deviations of
    the mixed run after the first 1000 steps (when nsteps >= 1000)."""
    import numpy as np

    from simson_amd import fields
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU

    def stats(u):
        mean, rms_, uv = profiles(u)
        return {"U": mean[0], "urms": rms_[0], "vrms": rms_[1], "wrms": rms_[2], "uv": uv,
                "tau_w": np.array([wall_shear(mean[0], re)])}

    f = fields.channel_field(nx, ny, nz, xl, zl, re, kmax=kmax, rms=rms, seed=seed)
    with BlaGPU.from_ufield(f) as g:
        g.upload_velocity(f.u, -2.0 / re)
        if spinup:
            g.run(spinup, dt, icfl=3)
        u0 = g.velocities()
        px0 = g.download(0)[3]

    def run_from(u, prec):
        acc, n, early = None, 0, None
        with BlaGPU.from_ufield(f, precision=prec) as g:
            g.upload_velocity(u, px0)
            for blk in range(nsteps // every):
                g.run(every, dt, icfl=3, first=(blk == 0))
                step = (blk + 1) * every
                s = stats(g.velocities())
                if step == 1000:
                    early = s
                if step > nsteps // 2 - every:
                    acc = s if acc is None else {k: acc[k] + s[k] for k in s}
                    n += 1
        return early, {k: v / n for k, v in acc.items()}

    runs = {}
    todo = [("fp64", SIMSON_FP64, None), ("mixed", SIMSON_MIXED, None)]
    todo += [(f"control{c}", SIMSON_FP64, c) for c in controls]
    for name, prec, cseed in todo:
        u = u0
        if cseed is not None:
            rng = np.random.default_rng(cseed)
            u = u0 * (1.0 + eps * rng.standard_normal(u0.shape))
        if say:
            say(f"    {name}: {nsteps} steps")
        runs[name] = run_from(u, prec)
    e64, a64 = runs["fp64"]
    emx, amx = runs["mixed"]
    out = {"averaged": {}, "steps": nsteps, "spinup": spinup, "every": every}
    for k in a64:
        dc = max(deviation(runs[f"control{c}"][1][k], a64[k]) for c in controls)
        out["averaged"][k] = {"mixed": deviation(amx[k], a64[k]), "control": dc}
    if e64 is not None:
        out["after_1000"] = {k: deviation(emx[k], e64[k]) for k in ("U", "urms", "vrms", "wrms")}
    return out
