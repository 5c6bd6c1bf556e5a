"""CPU restatement of the learned-closure network for tests (test
infrastructure only; the product path never imports it).

The same network as simson_amd/csrc/sg_mlp.hip: inputs padded to 16, two
hidden layers of 64 with exact GELU, outputs padded to 16, loss = mean
squared error over the real points and the six outputs.  ``bf16=True``
rounds every product operand to BF16 (round to nearest even) exactly where
the kernels do, with products accumulated in float64 here (FP32 there);
``bf16=False`` is the exact float64 network used to check the backward
formulas against finite differences.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf as _erf

NIN, NOUT, KI, NH, KO = 10, 6, 16, 64, 16
OW1 = 0
OB1 = OW1 + KI * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH * KO
NPAR = OB3 + KO

def bf16(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OW1:OB1].reshape(KI, NH), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH),
            p[OB2:OW3], p[OW3:OB3].reshape(NH, KO), p[OB3:NPAR])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward_backward(p, X, T, bf=True):
    """(loss, grad [NPAR], Y [n][6])."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = X.shape[0]
    Xp = np.zeros((n, KI))
    Xp[:, :NIN] = X
    Z1 = r(Xp) @ r(W1) + b1
    H1 = gelu(Z1)
    Z2 = r(H1) @ r(W2) + b2
    H2 = gelu(Z2)
    Y = r(H2) @ r(W3) + b3
    D = np.zeros((n, KO))
    D[:, :NOUT] = Y[:, :NOUT] - T
    loss = float(np.mean(D[:, :NOUT] ** 2))
    dY = D * (2.0 / (n * NOUT))
    dH2 = r(dY) @ r(W3).T
    dZ2 = dH2 * gelu_d(Z2)
    dH1 = r(dZ2) @ r(W2).T
    dZ1 = dH1 * gelu_d(Z1)
    g = np.zeros(NPAR)
    gW1 = r(Xp).T @ r(dZ1)
    gW1[NIN:] = 0.0
    gW3 = r(H2).T @ r(dY)
    gW3[:, NOUT:] = 0.0
    g[OW1:OB1] = gW1.ravel()
    g[OB1:OW2] = dZ1.sum(0)
    g[OW2:OB2] = (r(H1).T @ r(dZ2)).ravel()
    g[OB2:OW3] = dZ2.sum(0)
    g[OW3:OB3] = gW3.ravel()
    gb3 = dY.sum(0)
    gb3[NOUT:] = 0.0
    g[OB3:NPAR] = gb3
    return loss, g, Y[:, :NOUT]


def glorot(seed: int = 0) -> np.ndarray:
    """Random parameters in the kernel layout (padded rows/columns zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPAR)
    W1 = np.zeros((KI, NH))
    W1[:NIN] = rng.uniform(-1, 1, (NIN, NH)) * math.sqrt(6 / (NIN + NH))
    W3 = np.zeros((NH, KO))
    W3[:, :NOUT] = rng.uniform(-1, 1, (NH, NOUT)) * math.sqrt(6 / (NH + NOUT))
    p[OW1:OB1] = W1.ravel()
    p[OB1:OW2] = rng.uniform(-0.1, 0.1, NH)
    p[OW2:OB2] = (rng.uniform(-1, 1, (NH, NH)) * math.sqrt(6 / (2 * NH))).ravel()
    p[OB2:OW3] = rng.uniform(-0.1, 0.1, NH)
    p[OW3:OB3] = W3.ravel()
    p[OB3:OB3 + NOUT] = rng.uniform(-0.1, 0.1, NOUT)
    return p
