"""Learned subgrid closure: the network of the C ABI (simson_mlp_*) from Python.

``SgsNet`` wraps one device network (10 -> 64 -> 64 -> 6, GELU, BF16 matrix
cores with FP32 accumulation); ``Standardiser`` holds the input/output
scalings; ``train_closure`` trains on the a-priori samples of
``simson_amd.sgsdata`` and ``evaluate`` reports the a-priori errors against
the baselines (no model, Smagorinsky with a least-squares constant).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import gpu as _gpu
from . import sgsdata

NIN, NOUT = 10, 6
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_bound = False


def _lib():
    global _bound
    L = _gpu.lib()
    if not _bound:
        vp = C.c_void_p
        L.simson_mlp_create.argtypes = [C.c_ulonglong, C.POINTER(vp)]
        L.simson_mlp_nparams.argtypes = []
        L.simson_mlp_params.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.simson_mlp_set_data.argtypes = [vp, _fp, _fp, C.c_int]
        L.simson_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_void_p]
        L.simson_mlp_grad.argtypes = [vp, _fp, _fp, C.c_int, _fp, C.POINTER(C.c_float)]
        L.simson_mlp_infer.argtypes = [vp, _fp, C.c_int, _fp]
        L.simson_mlp_last_error.restype = C.c_char_p
        L.simson_mlp_last_error.argtypes = [vp]
        L.simson_mlp_destroy.argtypes = [vp]
        L.simson_mlp_destroy.restype = None
        for n in ("simson_mlp_create", "simson_mlp_nparams", "simson_mlp_params",
                  "simson_mlp_set_data", "simson_mlp_train", "simson_mlp_grad",
                  "simson_mlp_infer"):
            getattr(L, n).restype = C.c_int
        _bound = True
    return L


def nparams() -> int:
    return _lib().simson_mlp_nparams()


class SgsNet:
    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        st = _lib().simson_mlp_create(seed, C.byref(h))
        if st != 0:
            raise _gpu.SimsonError(st, "simson_mlp_create failed")
        self.h = h

    def _check(self, st):
        if st != 0:
            msg = _lib().simson_mlp_last_error(self.h)
            raise _gpu.SimsonError(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            _lib().simson_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(nparams(), np.float32)
        self._check(_lib().simson_mlp_params(self.h, p.ctypes.data_as(C.c_void_p), None))
        return p

    @params.setter
    def params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        self._check(_lib().simson_mlp_params(self.h, None, p.ctypes.data_as(C.c_void_p)))

    def set_data(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        self._check(_lib().simson_mlp_set_data(self.h, X, T, X.shape[0]))

    def train(self, steps: int, batch: int = 4096, lr: float = 1e-3) -> np.ndarray:
        losses = np.zeros(steps, np.float32)
        self._check(_lib().simson_mlp_train(self.h, steps, batch, lr,
                                            losses.ctypes.data_as(C.c_void_p)))
        return losses

    def grad(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(nparams(), np.float32)
        loss = C.c_float(0.0)
        self._check(_lib().simson_mlp_grad(self.h, X, T, X.shape[0], g, C.byref(loss)))
        return g, float(loss.value)

    def infer(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], NOUT), np.float32)
        self._check(_lib().simson_mlp_infer(self.h, X, X.shape[0], Y))
        return Y


@dataclass
class Standardiser:
    xm: np.ndarray
    xs: np.ndarray
    tm: np.ndarray
    ts: np.ndarray

    @classmethod
    def fit(cls, X, T):
        return cls(X.mean(0), X.std(0) + 1e-12, T.mean(0), T.std(0) + 1e-12)

    def x(self, X):
        return ((X - self.xm) / self.xs).astype(np.float32)

    def t(self, T):
        return ((T - self.tm) / self.ts).astype(np.float32)

    def t_inv(self, Tn):
        return Tn.astype(np.float64) * self.ts + self.tm


def train_closure(X, T, steps=3000, batch=4096, lr=2e-3, seed=1, shuffle_seed=7):
    """Train a network on a-priori samples (wall units); returns (net,
    standardiser, losses)."""
    sc = Standardiser.fit(X, T)
    rng = np.random.default_rng(shuffle_seed)
    idx = rng.permutation(X.shape[0])
    net = SgsNet(seed)
    net.set_data(sc.x(X[idx]), sc.t(T[idx]))
    losses = net.train(steps, batch, lr)
    return net, sc, losses


def deviatoric(T: np.ndarray) -> np.ndarray:
    """Remove one third of the trace from tau_11, tau_22, tau_33."""
    out = T.copy()
    tr = (T[:, 0] + T[:, 3] + T[:, 5]) / 3.0
    out[:, 0] -= tr
    out[:, 3] -= tr
    out[:, 5] -= tr
    return out


def evaluate(net: SgsNet, sc: Standardiser, X, T, delta2_plus: float):
    """A-priori errors on held-out samples: relative mean-square errors of
    the full stresses (network; no model is 1) and of the deviatoric
    stresses (network against the Smagorinsky model with the constant c_s^2
    fitted by least squares on these very samples, the most favourable
    static constant), and the correlations of the predicted tau_12."""
    P = sc.t_inv(net.infer(sc.x(X)))
    mse = lambda a: float(np.mean(a * a))
    out = {"rel_mse_full": mse(P - T) / mse(T)}
    Td, Pd = deviatoric(T), deviatoric(P)
    M = sgsdata.smagorinsky_basis(X, delta2_plus)
    cs2 = float(np.sum(M * Td) / np.sum(M * M))
    Sd = cs2 * M
    out["cs2_fitted"] = cs2
    out["rel_mse_dev_net"] = mse(Pd - Td) / mse(Td)
    out["rel_mse_dev_smag"] = mse(Sd - Td) / mse(Td)
    out["corr_tau12_net"] = float(np.corrcoef(P[:, 1], T[:, 1])[0, 1])
    out["corr_tau12_smag"] = float(np.corrcoef(Sd[:, 1], T[:, 1])[0, 1])
    return out


def closure_maps(sc: Standardiser, re: float, re_tau: float, ny: int):
    """Feature and output maps of simson_gpu_closure for a network trained
    on les_samples data (wall units of re_tau): (in_scale[9], in_shift[9],
    yfeat[ny], out_scale[6], out_shift[6])."""
    utau = re_tau / re
    in_scale = 1.0 / (utau * utau * re * sc.xs[:9])
    in_shift = -sc.xm[:9] / sc.xs[:9]
    y = np.cos(np.pi * np.arange(ny) / (ny - 1))
    yfeat = ((1.0 - np.abs(y)) * re_tau - sc.xm[9]) / sc.xs[9]
    out_scale = sc.ts * utau * utau
    out_shift = sc.tm * utau * utau
    return in_scale, in_shift, yfeat, out_scale, out_shift
