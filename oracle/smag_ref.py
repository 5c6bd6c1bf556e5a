"""CPU restatement of bla's eddy-viscosity LES models, sgs.i iles = 2 and
iles = 3 (test infrastructure).

Only the tests and the benchmark's CPU leg use this module, as the checker
of the GPU model.  It follows:

- preprbl.f:60-71   the filter width Delta^2 = (xl/nx zl/nz dy)^(2/3) per y;
- les.f:140-420     velocities and velocity gradients of a y plane on the
                    unpadded nx x nz grid (du/dy = dv/dx - om_z,
                    dv/dy = -du/dx - dw/dz, dw/dy = om_x + dv/dz), the z
                    oddball zeroed, u0low/w0low taken off the mean mode,
                    inverse transforms vcfftb (nz) and vrfftb (nx);
- les.f:622-700     S_ij and |S| = sqrt(2 S_ij S_ij);
- les.f:717-880     the dynamic procedure (Germano/Lilly, cs = 0): L_ij and
                    M_ij with the test filter filter_cutoff.f (2-D spectral
                    cutoff at pi/2: kx < nx/4, |kz| <= nz/4) and the test
                    width 4^(2/3) Delta^2;
- les.f:1380-1510   C(x, y) = 1/2 <M_ij L_ij>_z / <M_ij M_ij>_z with the
                    clipping ineg (1: C >= 0 and C = 0 on the wall planes;
                    2: C = 0 on the wall planes; 0: none), or C = cs^2;
- les.f:1530-1650   nu_t = C |S| Delta^2 (ineg = 2: nu_t >= -1/re),
                    tau_ij = -2 nu_t S_ij / (nx nz), forward transforms
                    vrfftf (nx) and vcfftf (nz), z oddball zeroed;
- sgs_stress.f:62-131  sigma_i = i alpha tau_i1 + d tau_i2/dy + i beta tau_i3
                    (the y derivative of the Chebyshev interpolant).

iles = 3 (the HPF model, les.f:180-215, 404-620, 883-1380): the strain and
|S| come from the high-pass filtered velocities and vorticities that
linearbl.f:1871-2066 stores in gur (u - lpf (u - H_y^N u) with the ADM-RT
filters of les_ref, H_y^N u = u - L u), no u0low/w0low shift; the dynamic
procedure takes the Leonard term from the unfiltered velocities as for
iles = 2, and its B_ij from the secondary high-pass filtered fields of
linearbl.f:2067-2256 (the x-z filter of init_filter called with nx/4 + 1
and nz/4 + 1 points at bla.f:1520-1543, the same y filter).

Tensor order of tau and S: 11 12 13 22 23 33 (les.f:627-633).
Pinned by tests/test_smag.py against the reference bla compiled with the
eddy-viscosity models (tests/golden/smag_32x33x32.npz).
"""
from __future__ import annotations

import math

import numpy as np

PAIRS = ((0, 0), (1, 0), (2, 0), (1, 1), (2, 1), (2, 2))   # les.f:717-729


def delta2(nx: int, ny: int, nz: int, xl: float, zl: float) -> np.ndarray:
    eta = np.cos(math.pi * np.arange(ny) / (ny - 1))
    d = np.empty(ny)
    d[0] = d[-1] = 1.0 - eta[1]
    d[1:-1] = np.sqrt((eta[2:] - eta[1:-1]) * (eta[1:-1] - eta[:-2]))
    return (xl / nx * zl / nz * d) ** (2.0 / 3.0)


def to_phys(f: np.ndarray, nx: int) -> np.ndarray:
    """[nz, ny, nx/2] spectra -> [nz, ny, nx] physical (vcfftb, vrfftb)."""
    nz = f.shape[0]
    g = np.fft.ifft(f, axis=0) * nz
    X = np.zeros(g.shape[:2] + (nx // 2 + 1,), dtype=complex)
    X[..., : g.shape[2]] = g
    return np.fft.irfft(X, n=nx, axis=2) * nx


def to_spec(p: np.ndarray) -> np.ndarray:
    """[nz, ny, nx] physical -> [nz, ny, nx/2] (vrfftf, vcfftf, unnormalised)."""
    nx = p.shape[2]
    X = np.fft.rfft(p, axis=2)[..., : nx // 2]
    return np.fft.fft(X, axis=0)


def cutoff_filter(p: np.ndarray) -> np.ndarray:
    """filter_cutoff.f on physical [nz, ny, nx]: keep kx < nx/4, |kz| <= nz/4."""
    nz, _, nx = p.shape
    X = np.fft.rfft(p, axis=2)
    X[..., nx // 4:] = 0.0
    X = np.fft.fft(X, axis=0)
    X[nz // 4 + 1: nz - nz // 4] = 0.0
    X = np.fft.ifft(X, axis=0) * nz
    return np.fft.irfft(X, n=nx, axis=2) * nx / (nx * nz)


def strain(v: np.ndarray, alfa: np.ndarray, beta: np.ndarray, nx: int):
    """S_ij (order 11 12 13 22 23 33) at physical x-z of the five spectral
    fields v [5][nz][ny][nx/2] (u, v, w, om_x, om_z, the z oddball zeroed):
    du/dy = dv/dx - om_z, dv/dy = -du/dx - dw/dz, dw/dy = om_x + dv/dz."""
    u, ox, oz = v[:3], v[3], v[4]
    ia = 1j * alfa[None, None, :]
    ib = 1j * beta[:, None, None]
    g = [[None] * 3 for _ in range(3)]
    for i in range(3):
        g[i][0] = ia * u[i]
        g[i][2] = ib * u[i]
    g[0][1] = ia * u[1] - oz
    g[1][1] = -ia * u[0] - ib * u[2]
    g[2][1] = ox + ib * u[1]
    gp = [[to_phys(g[i][j], nx) for j in range(3)] for i in range(3)]
    return [gp[0][0], .5 * (gp[1][0] + gp[0][1]), .5 * (gp[2][0] + gp[0][2]), gp[1][1],
            .5 * (gp[1][2] + gp[2][1]), gp[2][2]]


def hpf_filters(nx: int, ny: int, nz: int, cutoff: float, iord: int):
    """(lpf, lpf2, L) of the HPF model: the primary x-z transfer function and
    y low-pass of les_ref, and the secondary x-z transfer function, whose
    homogeneous filters are those of a grid with half the wavenumbers
    (zero above kx = nx/4 and |kz| = nz/4)."""
    import les_ref
    lpf, L = les_ref.filter_operators(nx, ny, nz, cutoff, iord)
    gx = np.zeros(nx // 2)
    gz = np.zeros(nz // 2 + 1)
    gx[:nx // 4 + 1] = les_ref.homogeneous_transfer(nx // 4 + 1, cutoff)
    gz[:nz // 4 + 1] = les_ref.homogeneous_transfer(nz // 4 + 1, cutoff)
    kz = np.array([z if z <= nz // 2 else nz - z for z in range(nz)])
    lpf2 = 1.0 - (1.0 - np.outer(gz[kz], gx)) ** (iord + 1)
    return lpf, lpf2, L


def les_tau(ur: np.ndarray, ui: np.ndarray, alfa: np.ndarray, beta: np.ndarray, nx: int,
            dlt2: np.ndarray, cs: float = 0.0, ineg: int = 1, re: float = 1.0,
            u0low: float = 0.0, w0low: float = 0.0, hpf=None) -> np.ndarray:
    """Main storage slots 0-4 (u, v, w, om_x, om_z at physical y) ->
    normalised stress spectra tau [6][nz][ny][nx/2]; hpf = (lpf, lpf2, L)
    of hpf_filters for iles = 3."""
    nz, ny = ur.shape[1], ur.shape[2]
    odd = nz // 2
    v = (ur[:5] + 1j * ui[:5]).copy()
    v[:, odd] = 0.0
    if hpf is None:
        v[0, 0, :, 0] -= u0low
        v[2, 0, :, 0] -= w0low
        vs = v
    else:
        lpf, lpf2, L = hpf
        W = np.einsum("yj,szjx->szyx", L, v)
        vs = v - lpf[None, :, None, :] * W
        vb = v - lpf2[None, :, None, :] * W
    up = [to_phys(v[i], nx) for i in range(3)]
    S = strain(vs, alfa, beta, nx)
    sabs = np.sqrt(2. * (S[0] ** 2 + S[3] ** 2 + S[5] ** 2) +
                   4. * (S[1] ** 2 + S[2] ** 2 + S[4] ** 2))
    d2 = dlt2[None, :, None]
    if cs == 0.0:
        dr = sabs * d2
        L0 = [cutoff_filter(-up[a] * up[b]) for a, b in PAIRS]
        M0 = [cutoff_filter(-S[k] * dr) for k in range(6)]
        SB = S if hpf is None else strain(vb, alfa, beta, nx)
        B = [cutoff_filter(SB[k]) for k in range(6)]
        uf = [cutoff_filter(up[i]) for i in range(3)]
        L = [L0[k] + uf[a] * uf[b] for k, (a, b) in enumerate(PAIRS)]
        tsabs = np.sqrt(2. * (B[0] ** 2 + B[3] ** 2 + B[5] ** 2) +
                        4. * (B[1] ** 2 + B[2] ** 2 + B[4] ** 2))
        drt = 4.0 ** (2.0 / 3.0) * d2 * tsabs
        M = [M0[k] + B[k] * drt for k in range(6)]
        w = (1., 2., 2., 1., 2., 1.)
        cnum = sum(w[k] * M[k] * L[k] for k in range(6)).sum(axis=0)   # [ny, nx]
        cden = sum(w[k] * M[k] * M[k] for k in range(6)).sum(axis=0)
        C = .5 * cnum / cden
        if ineg == 1:
            C = np.maximum(C, 0.0)
        if ineg in (1, 2):
            C[0] = 0.0
            C[-1] = 0.0
        C = np.broadcast_to(C[None], sabs.shape)
    else:
        C = np.full(sabs.shape, cs * cs)
    nut = C * sabs * d2
    if ineg == 2:
        nut = np.maximum(nut, -1.0 / re)
    gd2 = 1.0 / (nx * nz)
    tau = np.stack([to_spec(-2.0 * nut * gd2 * S[k]) for k in range(6)])
    tau[:, odd] = 0.0
    return tau


def cheb_dy(f: np.ndarray) -> np.ndarray:
    """d/dy along axis 1 of values on the Gauss-Lobatto points
    y_j = cos(pi j/(ny-1)): vchbf (scaled by 2/(ny-1)), dcheb, vchbb."""
    ny = f.shape[1]
    N = ny - 1
    j = np.arange(ny)
    T = np.cos(math.pi * np.outer(j, j) / N)           # T[k, j] = T_k(y_j)
    c = np.ones(ny)
    c[0] = c[-1] = 0.5
    A = (2.0 / N) * (c[:, None] * T * c[None, :])      # analysis
    a = np.einsum("kj,zjx->zkx", A, f)
    b = np.zeros_like(a)
    b[:, N - 1] = 2 * N * a[:, N]
    for i in range(N - 2, -1, -1):
        b[:, i] = b[:, i + 2] + 2 * (i + 1) * a[:, i + 1]
    b[:, 0] *= 0.5
    return np.einsum("jk,zkx->zjx", T.T, b)


def sgs_force(tau: np.ndarray, alfa: np.ndarray, beta: np.ndarray) -> np.ndarray:
    """sigma [3][nz][ny][nx/2] at physical y from tau [6][nz][ny][nx/2]."""
    ia = 1j * alfa[None, None, :]
    ib = 1j * beta[:, None, None]
    s = np.empty((3,) + tau.shape[1:], dtype=complex)
    s[0] = ia * tau[0] + ib * tau[2] + cheb_dy(tau[1])
    s[1] = ia * tau[1] + ib * tau[4] + cheb_dy(tau[3])
    s[2] = ia * tau[2] + ib * tau[5] + cheb_dy(tau[4])
    return s


def run(o, ur, ui, puw, nsteps: int, dt: float, model) -> None:
    """nsteps RK3 steps of the oracle o with the force model(ur, ui) ->
    sigma evaluated at the start of every substep (bla.f:1945-2020, before
    the nonlinbl loop); the weight bookkeeping of orc_run (bla.f:1465-1477,
    2384-2397)."""
    from pyoracle import lib
    anrk = (8. / 15., 5. / 12., 3. / 4.)
    bnrk = (0., -17. / 60., -5. / 12.)
    anp1, bnp1 = dt * anrk[0], 0.0
    o.prhs(ur, ui, puw, anp1)
    for it in range(1, 3 * nsteps + 1):
        sig = model(ur, ui)
        sr = np.ascontiguousarray(sig.real)
        si = np.ascontiguousarray(sig.imag)
        lib().orc_set_force(o.h, sr.ctypes.data, si.ctypes.data)
        o.nonlin(ur, ui, do_cfl=False)
        an, bn = anp1, bnp1
        anp1, bnp1 = dt * anrk[it % 3], dt * bnrk[it % 3]
        o.linear(ur, ui, puw, an, bn, anp1, bnp1)
    lib().orc_set_force(o.h, None, None)
