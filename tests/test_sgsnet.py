"""Learned subgrid closure: data, network kernels, training, a-priori gain.

Pins:
* the data pipeline (simson_amd.sgsdata) against closed forms: for the
  single mode u = cos(x) the filtered stress is tau_11 = (1 - G(1)^2)/2 +
  (G(2) - G(1)^2) cos(2x)/2 with the top-hat transfer function G, the other
  stresses vanish and du/dx = -G(1) sin(x); filtered gradients of a
  divergence-free field are divergence-free;
* the CPU network (oracle/mlp_ref.py) backward pass against central finite
  differences of its own loss (float64, no rounding), and its BF16 path
  against the exact one;
* on the GPU, the BF16 matrix-core kernels (sg_mlp.hip) against the CPU
  network with the same BF16 operand rounding: gradients and outputs to
  1e-2 relative (FP32 against float64 accumulation, and BF16 roundings of
  intermediate values that differ in the last FP32 bits), loss to 1e-4;
* the a-priori gain: trained on filtered DNS fields that this solver
  generates (128x97x128 channel, Re_cl 4200, Re_tau ~ 200), on a held-out
  snapshot the
  network's stresses beat no model and the Smagorinsky model (its constant
  fitted by least squares on the same snapshot).
"""
import math
import sys

import numpy as np
import pytest

import mlp_ref
from simson_amd import fields, sgsdata


def test_filtered_single_mode_closed_form():
    nx, ny, nz, xl, zl = 32, 9, 16, 2 * math.pi, 2 * math.pi
    u = np.zeros((3, nz, ny, nx // 2), complex)
    u[0, 0, :, 1] = 0.5                                # u = cos(x)
    g, tau = sgsdata.filtered(u, nx, ny, nz, xl, zl, fw=4)
    G = lambda k: np.sinc(k * 4 * (xl / nx) / 2 / np.pi)
    x = np.arange(nx // 4) * 4 * xl / nx
    expect = 0.5 * (1 - G(1) ** 2) + 0.5 * (G(2) - G(1) ** 2) * np.cos(2 * x)
    assert np.abs(tau[0] - expect[None, None, :]).max() < 1e-14
    assert np.abs(tau[1:]).max() < 1e-14
    assert np.abs(g[0] + G(1) * np.sin(x)[None, None, :]).max() < 1e-14
    assert np.abs(g[1:]).max() < 1e-14


def test_filtered_gradients_divergence_free():
    f = fields.channel_field(32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0, kmax=8,
                             rms=1e-2, seed=3)
    X, T, re_tau = sgsdata.les_samples(f.u, 32, 33, 32, 4 * math.pi, 2 * math.pi, 3000.0)
    assert X.shape == (33 * 8 * 8, 10) and T.shape == (33 * 8 * 8, 6)
    assert np.abs(X[:, 0] + X[:, 4] + X[:, 8]).max() < 1e-12 * np.abs(X[:, :9]).max()
    # diagonal stresses of a positive filter are non-negative on average
    assert T[:, 0].mean() > 0 and T[:, 3].mean() > 0 and T[:, 5].mean() > 0
    assert abs(re_tau - math.sqrt(2 * 3000.0)) < 1e-6 * re_tau   # Poiseuille mean


def test_oracle_backward_matches_finite_differences():
    p = mlp_ref.glorot(3)
    rng = np.random.default_rng(1)
    X, T = rng.standard_normal((37, 10)), rng.standard_normal((37, 6))
    _, g, _ = mlp_ref.forward_backward(p, X, T, bf=False)
    for i in (mlp_ref.OW1 + 5, mlp_ref.OB1 + 3, mlp_ref.OW2 + 100, mlp_ref.OB2 + 7,
              mlp_ref.OW3 + 2 * 16 + 4, mlp_ref.OB3 + 5):
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (mlp_ref.forward_backward(pp, X, T, bf=False)[0] -
              mlp_ref.forward_backward(pm, X, T, bf=False)[0]) / (2 * e)
        assert abs(g[i] - fd) <= 1e-5 * max(abs(fd), 1e-8), (i, g[i], fd)
    # padded input rows and output columns carry no gradient
    gW1 = g[mlp_ref.OW1:mlp_ref.OB1].reshape(16, 64)
    assert np.all(gW1[10:] == 0)
    _, gb, _ = mlp_ref.forward_backward(p, X, T, bf=True)
    assert np.linalg.norm(gb - g) / np.linalg.norm(g) < 2e-2


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_mlp_kernels_match_oracle():
    from simson_amd.sgsnet import SgsNet
    print("[gpu] BF16 MLP kernels vs CPU network", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(2)
    p = mlp_ref.glorot(5).astype(np.float32)
    for n in (1000, 37):                          # full tiles and a ragged one
        X = rng.standard_normal((n, 10)).astype(np.float32)
        T = rng.standard_normal((n, 6)).astype(np.float32)
        with SgsNet(1) as net:
            net.params = p
            assert np.array_equal(net.params, p)
            g, loss = net.grad(X, T)
            Y = net.infer(X)
        lref, gref, Yref = mlp_ref.forward_backward(p.astype(np.float64), X, T, bf=True)
        eg, ey = _rel(g, gref), _rel(Y, Yref)
        print(f"    n {n}: grad rel {eg:.2e}, output rel {ey:.2e}, loss {loss:.6f} vs {lref:.6f}",
              flush=True)
        assert eg < 1e-2 and ey < 1e-2 and abs(loss - lref) < 1e-4 * lref


@pytest.mark.gpu
def test_closure_a_priori_gain():
    """Filtered DNS fields of this solver (the seeded 10% perturbation of
    the Re_cl 4200 channel after 2000 steps is turbulent, Re_tau ~ 200):
    train on the snapshots after 2000, 2500 and 3000 steps, test on the one
    after 4000 steps."""
    from simson_amd.gpu import BlaGPU
    from simson_amd.sgsnet import evaluate, train_closure
    print("[gpu] learned closure a-priori: DNS 128x97x128, 4x filter", flush=True)
    sys.stdout.flush()
    nx, ny, nz, re = 128, 97, 128, 4200.0
    xl, zl = 4 * math.pi, 4 * math.pi / 3
    f = fields.channel_field(nx, ny, nz, xl, zl, re, kmax=16, rms=0.1, seed=2)
    snaps = []
    with BlaGPU(nx, ny, nz, re, xl, zl, cflux=True) as g:
        g.upload_velocity(f.u, -2.0 / re)
        done = 0
        for n in (2000, 2500, 3000, 4000):
            g.run(n - done, 0.005, icfl=3, first=done == 0)
            done = n
            snaps.append(g.velocities())
    data = [sgsdata.les_samples(u, nx, ny, nz, xl, zl, re, fw=4) for u in snaps]
    Xtr = np.concatenate([d[0] for d in data[:3]])
    Ttr = np.concatenate([d[1] for d in data[:3]])
    Xte, Tte, re_tau = data[3]
    net, sc, losses = train_closure(Xtr, Ttr, steps=3000, batch=4096, lr=2e-3)
    with net:
        # filter widths in x and z in wall units: delta^2 = dx_f dz_f
        delta2 = (4 * xl / nx * re_tau) * (4 * zl / nz * re_tau)
        m = evaluate(net, sc, Xte, Tte, delta2)
    print(f"    loss {losses[0]:.4f} -> {losses[-1]:.4f}; Re_tau {re_tau:.1f}", flush=True)
    print("    " + ", ".join(f"{k} {v:.3f}" for k, v in m.items()), flush=True)
    assert losses[-100:].mean() < 0.6 * losses[:10].mean()
    assert m["rel_mse_full"] < 0.5                    # no model: 1
    assert m["rel_mse_dev_net"] < m["rel_mse_dev_smag"]
    assert m["corr_tau12_net"] > max(0.5, m["corr_tau12_smag"])
