"""A-posteriori: LES runs on the coarse grid with no model, with ADM-RT (the
reference's own closure, the yardstick) and with the learned closure, against
the filtered DNS of this solver.

The DNS is the 128x97x128 channel (Re_cl 4200) from the seeded 10%
perturbation; after 2000 steps it is turbulent (Re_tau ~ 200).  The network
is trained a priori on its filtered snapshots after 2000, 2500 and 3000
steps (4x top-hat in x and z).  The LES grid is every fourth point in x and
z (32x97x32); the LES starts from the filtered DNS field after 3000 steps
and runs 1000 steps beside the DNS.  Reported: relative errors of the mean
profile and of the u, v, w rms profiles against the filtered DNS at the
end.  Measured (seeded, deterministic): mean profile 7.9e-3 / 9.8e-3 /
4.2e-3 and rms errors (u v w) 0.24 0.29 0.19 / 0.21 0.23 0.15 / 0.058 0.075
0.056 for no model / ADM-RT / learned.  Asserted: every run stays finite
and within 10% of the DNS mean profile, and the learned closure's rms
errors are below 3/4 of those without a model in every component.
"""
import math
import sys

import numpy as np
import pytest

from simson_amd import fields, sgsdata

NX, NY, NZ, RE, DT = 128, 97, 128, 4200.0, 0.005
XL, ZL = 4 * math.pi, 4 * math.pi / 3
FW = 4
LX, LZ = NX // FW, NZ // FW


def to_les(u):
    """Filtered DNS spectra restricted to the LES modes (|kx| < LX/2,
    |kz| < LZ/2), storage layout of the LES grid."""
    G = sgsdata.tophat(NX, NZ, XL, ZL, FW)[:, :NX // 2]           # [nz, nx/2]
    kz = fields.kz_of_index(NZ)
    out = np.zeros((3, LZ, NY, LX // 2), complex)
    for zl, k in enumerate(fields.kz_of_index(LZ)):
        if abs(k) >= LZ // 2:
            continue                                               # oddball of the LES grid
        zd = int(np.nonzero(kz == k)[0][0])
        out[:, zl, :, :] = u[:, zd, :, :LX // 2] * G[zd, None, :LX // 2]
    return out


def profiles(u, nz):
    """Mean U(y) and the rms of u, v, w over x and z (boxamp.f weights)."""
    w = np.full(u.shape[-1], 2.0)
    w[0] = 1.0
    a = np.abs(u) ** 2 * w
    a[:, 0, :, 0] = 0.0
    a[:, nz // 2] = 0.0
    return u[0, 0, :, 0].real, np.sqrt(a.sum(axis=(1, 3)))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_les_a_posteriori():
    from simson_amd.gpu import BlaGPU
    from simson_amd.sgsnet import closure_maps, train_closure
    print("[gpu] a-posteriori LES 32x97x32 against filtered DNS 128x97x128", flush=True)
    sys.stdout.flush()
    f = fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=16, rms=0.1, seed=2)
    snaps = {}
    with BlaGPU(NX, NY, NZ, RE, XL, ZL, cflux=True) as g:
        g.upload_velocity(f.u, -2.0 / RE)
        done = 0
        for n in (2000, 2500, 3000, 4000):
            g.run(n - done, DT, icfl=3, first=done == 0)
            done = n
            snaps[n] = g.velocities()
    data = [sgsdata.les_samples(snaps[n], NX, NY, NZ, XL, ZL, RE, fw=FW) for n in (2000, 2500, 3000)]
    X = np.concatenate([d[0] for d in data])
    T = np.concatenate([d[1] for d in data])
    re_tau = float(np.mean([d[2] for d in data]))
    net, sc, _ = train_closure(X, T, steps=3000, batch=4096, lr=2e-3)
    params = net.params
    net.close()
    Uref, Rref = profiles(to_les(snaps[4000]), LZ)
    u0 = to_les(snaps[3000])
    results = {}
    for model in ("none", "adm-rt", "learned"):
        with BlaGPU(LX, NY, LZ, RE, XL, ZL, cflux=True) as g:
            if model == "adm-rt":
                g.les(iles=1, cutoff=2.0 / 3.0, iord=5, chi=0.2)
            elif model == "learned":
                g.closure(params, *closure_maps(sc, RE, re_tau, NY))
            g.upload_velocity(u0, -2.0 / RE)
            g.run(1000, DT, icfl=3)
            u = g.velocities()
        assert np.all(np.isfinite(u)), model
        U, R = profiles(u, LZ)
        results[model] = (rel(U, Uref), [rel(R[c], Rref[c]) for c in range(3)])
        print(f"    {model:8s}: mean profile {results[model][0]:.3e}, "
              f"rms u v w {results[model][1][0]:.3f} {results[model][1][1]:.3f} "
              f"{results[model][1][2]:.3f}", flush=True)
    for model, (eu, _) in results.items():
        assert eu < 0.1, (model, eu)
    for c in range(3):
        assert results["learned"][1][c] < 0.75 * results["none"][1][c], (c, results)
