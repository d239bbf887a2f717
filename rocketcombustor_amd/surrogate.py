"""Learned chemistry (no reference counterpart): an FP32 MLP for the species
source term, trained on the GPU by the library's own kernels
(``rc_mlp_*``), and its use in the steady step with a per-cell gate and the
exact FP64 kinetics as fallback (``rc_set_surrogate``).

Network: (T, p, Y_1..Y_Nsp) -> omega_1..omega_{Nsp-1}, three tanh hidden
layers of 128 (the configuration of the learned-chemistry workload).
Inputs are standardised (log p for the pressure); the outputs stand for
omega_k = knee s_k sinh(y_k) with s_k the largest |omega_k| of the data (a
squared-error fit then weighs relative errors); both transforms travel with
the network.

Training data are states the kinetics itself produces: snapshots of the
primitive field of steady runs with the production rates the device forms
for them (``SteadyCase.derived``).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Optional

import numpy as np

from . import native
from .native import NVAR, check, lib

N_IN = 6           # T, p, Y_1..Y_4
N_OUT = 3          # omega_1..omega_3


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """Device-resident network (``rc_mlp``)."""

    def __init__(self, n_in: int = N_IN, width: int = 128, hidden: int = 3, n_out: int = N_OUT, device: int = 0,
                 seed: int = 5):
        h = C.c_void_p()
        rc = lib().rc_mlp_create(int(device), n_in, width, hidden, n_out, C.byref(h))
        if rc:
            raise native.RcError(rc, lib().rc_mlp_last_error(None).decode())
        self.h = h
        self.n_in, self.n_out, self.width, self.hidden = n_in, n_out, width, hidden
        dims = [n_in] + [width] * hidden + [n_out]
        self.shapes = [(dims[k], dims[k + 1]) for k in range(len(dims) - 1)]
        self.nparam = int(lib().rc_mlp_num_params(h))
        self.params = glorot(self.shapes, seed)

    def _check(self, rc):
        if rc:
            raise native.RcError(rc, lib().rc_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().rc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.empty(self.nparam, dtype=np.float32)
        self._check(lib().rc_mlp_get_params(self.h, _fp(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._check(lib().rc_mlp_set_params(self.h, _fp(p)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((len(x), self.n_out), dtype=np.float32)
        self._check(lib().rc_mlp_forward(self.h, _fp(x), len(x), _fp(y)))
        return y

    def gradient(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(lib().rc_mlp_gradient(self.h, _fp(x), _fp(t), len(x), _fp(g), _fp(loss)))
        return float(loss[0]), g

    def set_data(self, x: np.ndarray, t: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._check(lib().rc_mlp_set_data(self.h, _fp(x), _fp(t), len(x)))

    def train(self, nsteps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 5, every: int = 100) -> np.ndarray:
        # the loss is recorded after every ``every`` steps and after the last
        nrep = (nsteps // every + (1 if nsteps % every else 0)) if every > 0 else 0
        out = np.zeros(max(nrep, 1), dtype=np.float32)
        self._check(lib().rc_mlp_train(self.h, int(nsteps), int(batch), float(lr), int(seed), int(every), _fp(out)))
        return out[:nrep]


def glorot(shapes, seed: int) -> np.ndarray:
    """Glorot-uniform weights, zero biases (seeded NumPy generator)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fi, fo in shapes:
        a = np.sqrt(6.0 / (fi + fo))
        parts += [rng.uniform(-a, a, size=fi * fo), np.zeros(fo)]
    return np.concatenate(parts).astype(np.float32)


@dataclasses.dataclass
class Transform:
    """Input standardisation (log p) and the output map: the network's y_k
    stands for omega_k = knee s_k sinh(y_k) (linear below knee s_k,
    logarithmic above, so the squared error weighs relative errors)."""
    mean: np.ndarray          # (N_IN,)
    std: np.ndarray           # (N_IN,)
    out_scale: np.ndarray     # (N_OUT,) s_k: the largest |omega_k| of the data
    knee: float
    lo: np.ndarray            # (N_IN,) training range of the raw inputs (T, p, Y)
    hi: np.ndarray

    @staticmethod
    def fit(raw: np.ndarray, omega: np.ndarray, knee: float = 1e-3) -> "Transform":
        z = features(raw)
        return Transform(mean=z.mean(axis=0), std=np.maximum(z.std(axis=0), 1e-12),
                         out_scale=np.maximum(np.abs(omega).max(axis=0), 1e-30), knee=float(knee),
                         lo=raw.min(axis=0), hi=raw.max(axis=0))

    def x(self, raw: np.ndarray) -> np.ndarray:
        return ((features(raw) - self.mean) / self.std).astype(np.float32)

    def y(self, omega: np.ndarray) -> np.ndarray:
        return np.arcsinh(omega / (self.knee * self.out_scale)).astype(np.float32)

    def omega(self, y: np.ndarray) -> np.ndarray:
        """Rates of network outputs y (float64, as the device maps them)."""
        return self.knee * self.out_scale * np.sinh(np.asarray(y, dtype=np.float64))

    def packed(self) -> np.ndarray:
        """The layout rc_set_surrogate reads: mean, std, out_scale, knee, lo, hi."""
        return np.concatenate([self.mean, self.std, self.out_scale, [self.knee], self.lo,
                               self.hi]).astype(np.float64)


def train_schedule(m: "MLP", nsteps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 5,
                   phases: int = 4, decay: float = 0.3) -> np.ndarray:
    """Adam over ``phases`` equal parts of ``nsteps`` with the learning rate
    multiplied by ``decay`` from one part to the next; returns the batch loss
    at the end of each part."""
    out = []
    per = max(nsteps // phases, 1)
    for k in range(phases):
        out.append(m.train(per, batch=batch, lr=lr * decay ** k, seed=seed + k, every=per)[-1])
    return np.array(out)


def gate(tf: Transform, raw: np.ndarray, w_net: np.ndarray, w_exact: np.ndarray, tol: float, rtol: float,
         trust: Optional["TrustRegion"] = None):
    """The per-cell acceptance the device gate forms (training range, mass
    balance within tol, with a trust region its marked grid cells, and with
    rtol > 0 agreement with the exact rates)."""
    inside = np.all((raw >= tf.lo) & (raw <= tf.hi), axis=1)
    ok = inside & (np.abs(w_net.sum(axis=1)) <= tol * (np.abs(w_net).sum(axis=1) + 1e-3 * tf.out_scale.max()))
    if trust is not None:
        idx = trust.index(raw)
        ok &= (idx >= 0) & (trust.accept[np.maximum(idx, 0)] != 0)
    if rtol > 0:
        ok &= np.all(np.abs(w_net - w_exact) <= rtol * np.maximum(np.abs(w_exact), 1e-3 * tf.out_scale), axis=1)
    return ok


@dataclasses.dataclass
class TrustRegion:
    """A calibrated trust region of a trained network, for a gate that needs
    no exact rate: a grid over the features (T, log p, Y_1) between lo and
    hi; a grid cell is accepted when the network's relative error on the
    held-out states falling in it, max_k |w_k - w_k,exact| / max(|w_k,exact|,
    1e-3 s_k), stayed within ``tau`` for at least ``nmin`` states."""
    dims: tuple
    lo: np.ndarray
    hi: np.ndarray
    accept: np.ndarray        # uint8, dims[0]*dims[1]*dims[2]
    tau: float
    nmin: int

    @staticmethod
    def feats(raw: np.ndarray) -> np.ndarray:
        return np.column_stack([raw[:, 0], np.log(raw[:, 1]), raw[:, 2]])

    def index(self, raw: np.ndarray):
        """Grid cell of each state (-1 outside the grid)."""
        f = self.feats(raw)
        u = (f - self.lo) * (np.array(self.dims) / (self.hi - self.lo))
        inside = np.all((u >= 0) & (u < np.array(self.dims)), axis=1)
        b = np.clip(np.floor(u).astype(np.int64), 0, np.array(self.dims) - 1)
        idx = (b[:, 0] * self.dims[1] + b[:, 1]) * self.dims[2] + b[:, 2]
        return np.where(inside, idx, -1)

    @staticmethod
    def calibrate(tf: "Transform", raw: np.ndarray, w_net: np.ndarray, w_exact: np.ndarray,
                  dims=(48, 16, 48), tau: float = 2.5e-4, nmin: int = 10) -> "TrustRegion":
        lo = np.array([tf.lo[0], np.log(tf.lo[1]), tf.lo[2]])
        hi = np.array([tf.hi[0], np.log(tf.hi[1]), tf.hi[2]])
        hi = np.where(hi > lo, hi, lo + 1.0)
        hi = hi + 1e-9 * (hi - lo)           # the largest states fall inside the last grid cell
        tr = TrustRegion(dims=tuple(int(d) for d in dims), lo=lo, hi=hi,
                         accept=np.zeros(int(np.prod(dims)), dtype=np.uint8), tau=float(tau), nmin=int(nmin))
        err = rel_error(tf, w_net, w_exact)
        idx = tr.index(raw)
        keep = idx >= 0
        n = np.bincount(idx[keep], minlength=tr.accept.size)
        worst = np.zeros(tr.accept.size)
        np.maximum.at(worst, idx[keep], err[keep])
        tr.accept[:] = ((n >= nmin) & (worst <= tau)).astype(np.uint8)
        return tr


def rel_error(tf: "Transform", w: np.ndarray, w_exact: np.ndarray) -> np.ndarray:
    """Per state: max_k |w_k - w_k,exact| / max(|w_k,exact|, 1e-3 s_k)."""
    return np.max(np.abs(w - w_exact) / np.maximum(np.abs(w_exact), 1e-3 * tf.out_scale), axis=1)


def features(raw: np.ndarray) -> np.ndarray:
    """(T, p, Y_1..Y_4) -> (T, log p, Y_1..Y_4)."""
    z = np.array(raw, dtype=np.float64, copy=True)
    z[:, 1] = np.log(z[:, 1])
    return z


def raw_inputs(pa: np.ndarray) -> np.ndarray:
    """(T, p, Y_1..Y_4) of primitive states PA(Nvar) (Y_4 = 1 - sum)."""
    y = pa[:, 4:NVAR]
    return np.column_stack([pa[:, 3], pa[:, 0], y, 1.0 - y.sum(axis=1)])


def sample_states(cases, steps: List[int], per_snapshot: int, seed: int = 5, device: int = 0):
    """States the kinetics produces: for each case, steady steps on the
    device with snapshots after each count in ``steps``; each snapshot
    contributes ``per_snapshot`` random cells (inputs and the device's exact
    FP64 production rates)."""
    from .solver import SteadyCase
    rng = np.random.default_rng(seed)
    xs, ws = [], []
    for case in cases:
        with SteadyCase(case, device=device) as g:
            done = 0
            for s in steps:
                if s > done:
                    g.run_steady(s - done, history=False)
                    done = s
                pa = g.primitive()
                w = g.derived()["production_rate"]
                idx = rng.choice(len(pa), size=min(per_snapshot, len(pa)), replace=False)
                xs.append(raw_inputs(pa[idx]))
                ws.append(w[idx])
    return np.concatenate(xs), np.concatenate(ws)
