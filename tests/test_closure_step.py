"""The learned closure inside the time step (simson_gpu_closure).

At every substep the library synthesises the nine velocity gradients on the
dealiased x-z grid, evaluates the network at every point on the BF16 matrix
cores, transforms the six stresses back and subtracts their divergence
sigma_i = i alpha tau_i1 + d tau_i2/dy + i beta tau_i3 from H_i
(sgs_stress.f:62-131, nonlinbl.f:1011-1050).

Pins (GPU):
* zero stresses leave the step bit-identical;
* spatially constant stresses have no divergence: the step moves by
  rounding only;
* the force of a random network on a random channel field equals the force
  computed on the CPU from the same definitions (gradients by spectral and
  Chebyshev differentiation on the 3/2-padded grid, the CPU network of
  oracle/mlp_ref.py with the same BF16 operand rounding, forward transform,
  divergence) to 1e-3 relative (BF16 roundings of features that land on
  the other side of a rounding boundary; FP32 network arithmetic).
"""
import math
import sys

import numpy as np
import pytest

import mlp_ref
from simson_amd import fields, sgsdata

NX, NY, NZ, RE = 64, 65, 48, 4200.0
XL, ZL = 4 * math.pi, 4 * math.pi / 3
ODD = NZ // 2


def _field():
    return fields.channel_field(NX, NY, NZ, XL, ZL, RE, kmax=12, rms=5e-2, seed=4)


def _maps(u):
    from simson_amd.sgsnet import Standardiser, closure_maps
    X, T, re_tau = sgsdata.les_samples(u, NX, NY, NZ, XL, ZL, RE, fw=4)
    sc = Standardiser.fit(X, T)
    return closure_maps(sc, RE, re_tau, NY)


def cpu_force(U5, params, in_scale, in_shift, yfeat, out_scale, out_shift):
    """-Delta H: the unnormalised divergence of the network's stresses,
    storage layout [3][nz][ny][nx/2]."""
    nxp, nzp = 3 * NX // 2, 3 * NZ // 2
    U5 = np.array(U5)
    U5[:, ODD] = 0.0
    u, v, w, ox, oz = U5
    al = (2 * np.pi / XL * np.arange(NX // 2))[None, None, :]
    be = (2 * np.pi / ZL * fields.kz_of_index(NZ))[:, None, None]
    ia, ib = 1j * al, 1j * be
    G = [ia * u, ia * v - oz, ib * u, ia * v, -(ia * u + ib * w), ib * v, ia * w, ox + ib * v,
         ib * w]
    phys = sgsdata._synth(sgsdata._full_spectrum(np.stack(G), NX), nxp, nzp, NZ)  # [9, ny, nzp, nxp]
    X = np.empty((NY * nzp * nxp, 10))
    for k in range(9):
        X[:, k] = (phys[k] * in_scale[k] + in_shift[k]).ravel()
    X[:, 9] = np.repeat(yfeat, nzp * nxp)
    # the device rounds features to FP32 first, then to BF16
    X = X.astype(np.float32).astype(np.float64)
    _, _, Y = mlp_ref.forward_backward(params.astype(np.float64), X,
                                       np.zeros((X.shape[0], 6)), bf=True)
    tau = (Y * out_scale + out_shift).T.reshape(6, NY, nzp, nxp)
    th = sgsdata._analyse(tau, NX, NZ)[..., :NX // 2] * (nxp * nzp)      # [6, ny, nz, nxh]
    th = np.swapaxes(th, 1, 2)                                              # [6, nz, ny, nxh]
    D = sgsdata.cheb_diff(NY)
    dy = lambda f: np.einsum("ab,zbx->zax", D, f)
    t11, t12, t13, t22, t23, t33 = th
    s1 = ia * t11 + dy(t12) + ib * t13
    s2 = ia * t12 + dy(t22) + ib * t23
    s3 = ia * t13 + dy(t23) + ib * t33
    return np.stack([s1, s2, s3])


def _nonlin_pair(form, closure_args):
    """H from one nonlin call without and with the closure, same state."""
    from simson_amd.gpu import SIMSON_FP64, SIMSON_MIXED, BlaGPU
    f = _field()
    prec = SIMSON_MIXED if form == "mixed" else SIMSON_FP64
    with BlaGPU.from_ufield(f, precision=prec) as g:
        g.upload_velocity(f.u, -2.0 / RE)
        g.prhs(8.0 / 15.0 * 0.005)
        ur, ui, puw, px = g.download(7)
        g.nonlin(False)
        h0r, h0i, _, _ = g.download(3)
        g.upload_state(ur, ui, puw, px)
        g.closure(*closure_args)
        g.nonlin(False)
        h1r, h1i, _, _ = g.download(3)
    U5 = ur[:5] + 1j * ui[:5]
    return U5, (h0r + 1j * h0i), (h1r + 1j * h1i)


def _masked(a):
    return np.delete(a, ODD, axis=1)


@pytest.mark.gpu
def test_zero_and_constant_stresses():
    print("[gpu] closure with zero and constant stresses", flush=True)
    sys.stdout.flush()
    f = _field()
    ins, insh, yf, _, _ = _maps(f.u)
    zero = np.zeros(mlp_ref.NPAR, np.float32)
    _, h0, h1 = _nonlin_pair("fp64", (zero, ins, insh, yf, np.zeros(6), np.zeros(6)))
    assert np.array_equal(h0, h1)
    _, h0, h1 = _nonlin_pair("fp64", (zero, ins, insh, yf, np.zeros(6),
                                      np.array([1e-3, 2e-4, -3e-4, 5e-4, 1e-4, 7e-4])))
    d = np.linalg.norm(_masked(h1 - h0)) / np.linalg.norm(_masked(h0))
    print(f"    constant stresses: relative change of H {d:.2e}", flush=True)
    assert d < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_closure_force_matches_cpu(form):
    print(f"[gpu] closure force vs CPU ({form}) {NX}x{NY}x{NZ}", flush=True)
    sys.stdout.flush()
    f = _field()
    maps = _maps(f.u)
    params = mlp_ref.glorot(11).astype(np.float32)
    U5, h0, h1 = _nonlin_pair(form, (params,) + maps)
    ref = cpu_force(U5, params, *maps)
    got = -(h1 - h0)
    e = np.linalg.norm(_masked(got - ref)) / np.linalg.norm(_masked(ref))
    scale = np.linalg.norm(_masked(ref)) / np.linalg.norm(_masked(h0))
    print(f"    force rel L2 {e:.2e} (force/H {scale:.2e})", flush=True)
    assert scale > 1e-3                      # the force is not negligible
    assert e < (1e-2 if form == "mixed" else 1e-3)
