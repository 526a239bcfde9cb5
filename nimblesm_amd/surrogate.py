"""Learned control variate for ensemble means of the reaction-force history.

A small network (nsm_mlp_* of the C ABI: BF16 weights and activations, FP32
accumulation on the matrix cores, trained on the device by Adam) maps a
sample's parameters (E, nu, rho, v_top) to its 64-point reaction history.
The ensemble mean over N draws is then estimated as

    mean_hat = mean_i f(x_i)  +  mean_{j in P} (y_j - f(x_j)),

where P is a random fraction of the draws solved by the plain solver: the
second term removes the network's bias.  The same plain solves, held out
from training, judge the estimate: the standard error of the correction
term (with the finite-population factor sqrt(1 - |P|/N)), relative to the
norm of the estimate, is the held-out estimate of its error; when it
exceeds the tolerance the estimator falls back to plain solves of every
draw.  (The network's own per-draw error on the held-out draws, limited by
its BF16 activations to a few 1e-3, is reported beside it; the bound applies
to the estimated mean.)

The reference has no learned part.  Its ensemble estimator is the linear
closure of its UQ model (UqModel::ApplyClosure, src/uq/nimble_uq.cc:327-345:
off-nominal forces as linear combinations of exact samples); the linear
control variate below (a least-squares linear map from the same features)
is the estimator the network is compared with at equal cost.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

from ._capi import check, lib
from .ensemble import NOMINAL, Draws

WIDTHS = (4, 256, 256, 256, 64)


class MLP:
    """Device network (see include/nimblesm_amd.h, nsm_mlp_*)."""

    def __init__(self, widths: Sequence[int] = WIDTHS, seed: int = 0):
        self.widths = tuple(int(w) for w in widths)
        w = np.array(self.widths, np.int32)
        h = C.c_void_p()
        check(lib().nsm_mlp_create(len(w) - 1, w.ctypes.data, C.c_uint64(seed), C.byref(h)))
        self._h = h
        n = C.c_int64()
        check(lib().nsm_mlp_num_parameters(self._h, C.byref(n)))
        self.num_parameters = n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().nsm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_parameters(self) -> np.ndarray:
        out = np.zeros(self.num_parameters, np.float32)
        check(lib().nsm_mlp_get_parameters(self._h, out.ctypes.data))
        return out

    def set_parameters(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_parameters
        check(lib().nsm_mlp_set_parameters(self._h, p.ctypes.data))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.widths[0])
        y = np.zeros((x.shape[0], self.widths[-1]), np.float32)
        check(lib().nsm_mlp_forward(self._h, x.shape[0], x.ctypes.data, y.ctypes.data))
        return y

    def loss_gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.widths[0])
        t = np.ascontiguousarray(t, np.float32).reshape(x.shape[0], self.widths[-1])
        g = np.zeros(self.num_parameters, np.float32)
        loss = C.c_double()
        check(lib().nsm_mlp_loss_gradients(self._h, x.shape[0], x.ctypes.data, t.ctypes.data, C.byref(loss),
                                           g.ctypes.data))
        return loss.value, g

    def train(self, x, t, iterations: int, lr: float = 2e-3, lr_final: float = 2e-5) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.widths[0])
        t = np.ascontiguousarray(t, np.float32).reshape(x.shape[0], self.widths[-1])
        hist = np.zeros(iterations)
        check(lib().nsm_mlp_train(self._h, x.shape[0], x.ctypes.data, t.ctypes.data, int(iterations), float(lr),
                                  float(lr_final), hist.ctypes.data))
        return hist


def features(d: Draws, nominal=NOMINAL) -> np.ndarray:
    """Standardised inputs: log-ratios of E and rho to nominal over 0.1, the
    Poisson ratio about nominal over 0.03, v_top about 1.25 over 0.43."""
    return np.stack([np.log(d.E / nominal[0]) / 0.1, (d.nu - nominal[1]) / 0.03, np.log(d.rho / nominal[2]) / 0.1,
                     (d.v_top - 1.25) / 0.433], axis=1)


@dataclass
class Scaler:
    mean: np.ndarray
    std: np.ndarray

    @classmethod
    def fit(cls, y):
        s = y.std(axis=0)
        return cls(y.mean(axis=0), np.where(s > 0, s, 1.0))

    def to(self, y):
        return (y - self.mean) / self.std

    def back(self, z):
        return z * self.std + self.mean


class LinearSurrogate:
    """Least-squares affine map from the features to the history."""

    def fit(self, d: Draws, y: np.ndarray):
        A = np.concatenate([np.ones((len(d), 1)), features(d)], axis=1)
        self.coef, *_ = np.linalg.lstsq(A, y, rcond=None)
        return self

    def predict(self, d: Draws) -> np.ndarray:
        return np.concatenate([np.ones((len(d), 1)), features(d)], axis=1) @ self.coef


class MLPSurrogate:
    """The device network trained on (features, standardised histories)."""

    def __init__(self, widths=WIDTHS, seed: int = 0):
        self.net = MLP(widths, seed)

    def fit(self, d: Draws, y: np.ndarray, iterations: int = 3000, lr: float = 2e-3, lr_final: float = 2e-5):
        self.scaler = Scaler.fit(y)
        self.loss = self.net.train(features(d), self.scaler.to(y), iterations, lr, lr_final)
        return self

    def predict(self, d: Draws) -> np.ndarray:
        return self.scaler.back(self.net.forward(features(d)).astype(np.float64))

    def close(self):
        self.net.close()


@dataclass
class Estimate:
    mean: np.ndarray              # estimated mean history [64]
    est_rel_err: float            # standard error of the correction / ||mean||: the judged error
    heldout_rel_err: float        # ||y - f(x)|| / ||y|| over the plain-solved (held-out) draws
    heldout_mean_rel_err: float   # ||mean(y) - mean(f(x))|| / ||mean(y)|| over those draws
    num_plain: int
    fallback: bool


def control_variate(surrogate, d: Draws, plain: Callable[[Draws], np.ndarray], fraction: float = 0.05,
                    tol: float = 1e-3, seed: int = 0, pred: Optional[np.ndarray] = None,
                    plain_idx: Optional[np.ndarray] = None) -> Estimate:
    """Control-variate estimate of the mean history over the draws d, with
    round(fraction * N) randomly chosen draws solved by `plain` (which
    returns their histories); falls back to plain solves of every draw when
    the judged error exceeds tol."""
    N = len(d)
    f = surrogate.predict(d) if pred is None else pred
    if plain_idx is None:
        rng = np.random.default_rng(seed)
        plain_idx = np.sort(rng.choice(N, size=max(2, int(round(fraction * N))), replace=False))
    y = plain(d.subset(plain_idx))
    r = y - f[plain_idx]
    n = len(plain_idx)
    mean = f.mean(axis=0) + r.mean(axis=0)
    se = r.std(axis=0, ddof=1) / np.sqrt(n) * np.sqrt(max(1.0 - n / N, 0.0))
    est = float(np.linalg.norm(se) / np.linalg.norm(mean))
    held = float(np.linalg.norm(r) / np.linalg.norm(y))
    held_mean = float(np.linalg.norm(r.mean(axis=0)) / np.linalg.norm(y.mean(axis=0)))
    if est > tol:
        return Estimate(plain(d).mean(axis=0), est, held, held_mean, N, True)
    return Estimate(mean, est, held, held_mean, n, False)
