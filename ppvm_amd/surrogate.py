"""Learned surrogate for parameter sweeps of the lattice TFIM.

An ensemble of MLPs (inputs: h, t, steps and sin/cos features; hidden layers
256-256-256 with GELU; scalar output; BF16 weights and activations with FP32
accumulation on MFMA) is trained on the GPU, by the library's own forward and
backward kernels, on expectation values produced by the GPU propagator.  At
query time the ensemble mean is used where the ensemble agrees (standard
deviation <= ``fallback_std``); elsewhere the propagator is run (one
propagation per distinct h, prefix reuse over t).
"""

from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _native
from .sweep import DT

_f32p = ctypes.POINTER(ctypes.c_float)
_declared = False


def _lib():
    global _declared
    L = _native.lib()
    if not _declared:
        L.ppvm_mlp_last_error.restype = ctypes.c_char_p
        L.ppvm_mlp_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
        L.ppvm_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.ppvm_mlp_param_count.restype = ctypes.c_uint64
        L.ppvm_mlp_param_count.argtypes = [ctypes.c_void_p]
        L.ppvm_mlp_get_params.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _f32p, ctypes.c_size_t]
        L.ppvm_mlp_set_params.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _f32p, ctypes.c_size_t]
        L.ppvm_mlp_train.argtypes = [ctypes.c_void_p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_float, ctypes.c_uint32, _f32p]
        L.ppvm_mlp_predict.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, _f32p, _f32p, _f32p]
        L.ppvm_mlp_grads.argtypes = [ctypes.c_void_p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p,
                                     ctypes.c_size_t, _f32p]
        L.ppvm_mlp_gemm_check.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p]
        _declared = True
    return L


def _check(rc):
    if rc != 0:
        msg = _lib().ppvm_mlp_last_error().decode("utf-8", "replace")
        if rc == _native.PPVM_EINVAL:
            raise ValueError(msg)
        raise _native.PpvmError(rc, msg)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_f32p)


def features(h, t, dt: float = DT) -> np.ndarray:
    """Network inputs for sweep points (h, t): scaled h, time and step count,
    and sin/cos of h*t_s at harmonics 1, 2, 4 and of pi*t_s.  Time enters as
    t_s = steps*dt, the time the Trotter circuit actually reaches
    (steps = round(t/dt)): grid points with equal step counts are the same
    physical point and get the same inputs."""
    h = np.asarray(h, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    steps = np.rint(t / dt)
    ts = steps * dt
    cols = [(h - 2.25) / 1.75, ts - 1.0, steps / 20.0 - 1.0]
    for k in (1.0, 2.0, 4.0):
        cols += [np.sin(k * h * ts), np.cos(k * h * ts)]
    cols += [np.sin(np.pi * ts), np.cos(np.pi * ts)]
    return np.stack(cols, axis=-1).astype(np.float32)


class MLPEnsemble:
    """Owning wrapper of a ``ppvm_mlp_handle``."""

    def __init__(self, in_dim: int, hidden: int = 256, members: int = 4, max_batch: int = 4096,
                 seed: int = 31):
        h = ctypes.c_void_p()
        _check(_lib().ppvm_mlp_create(in_dim, hidden, members, max_batch, seed, ctypes.byref(h)))
        self.h = h
        self.in_dim, self.hidden, self.members, self.max_batch = in_dim, hidden, members, max_batch

    def close(self):
        if getattr(self, "h", None) is not None:
            _lib().ppvm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def n_params(self) -> int:
        return int(_lib().ppvm_mlp_param_count(self.h))

    def params(self, member: int) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(_lib().ppvm_mlp_get_params(self.h, member, out.ctypes.data_as(_f32p), len(out)))
        return out

    def set_params(self, member: int, p: np.ndarray):
        p, pp = _f32(p)
        _check(_lib().ppvm_mlp_set_params(self.h, member, pp, len(p)))

    def train(self, X, Y, steps: int = 3000, lr: float = 2e-3, use_graph: bool = True) -> np.ndarray:
        X, xp = _f32(X)
        Y, yp = _f32(Y)
        loss = np.zeros(self.members, dtype=np.float32)
        _check(_lib().ppvm_mlp_train(self.h, xp, yp, len(Y), steps, lr, 1 if use_graph else 0,
                                     loss.ctypes.data_as(_f32p)))
        return loss

    def predict(self, X, per_member: bool = False):
        X, xp = _f32(X)
        n = len(X)
        mean = np.zeros(n, dtype=np.float32)
        sd = np.zeros(n, dtype=np.float32)
        pm = np.zeros((self.members, n), dtype=np.float32) if per_member else None
        _check(_lib().ppvm_mlp_predict(self.h, xp, n, mean.ctypes.data_as(_f32p), sd.ctypes.data_as(_f32p),
                                       pm.ctypes.data_as(_f32p) if per_member else None))
        return (mean, sd, pm) if per_member else (mean, sd)

    def grads(self, X, Y, member: int = 0):
        X, xp = _f32(X)
        Y, yp = _f32(Y)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(_lib().ppvm_mlp_grads(self.h, xp, yp, len(Y), member, g.ctypes.data_as(_f32p), len(g),
                                     loss.ctypes.data_as(_f32p)))
        return float(loss[0]), g


def gemm_check(A, B) -> np.ndarray:
    A, ap = _f32(A)
    B, bp = _f32(B)
    M, K = A.shape
    N = B.shape[0]
    C = np.zeros((M, N), dtype=np.float32)
    _check(_lib().ppvm_mlp_gemm_check(ap, bp, M, N, K, C.ctypes.data_as(_f32p)))
    return C


GRID_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "lattice_sweep_grid.npz")


def load_grid(path: str = GRID_FILE) -> dict:
    """Propagator grid of the lattice TFIM sweep written by ``python -m
    ppvm_amd.sweep --out FILE``: h (64), t (64), dt, grid[h, m] = overlap after
    m backward Trotter steps (m = 0..40), seconds[h] = propagation wall time per
    h value on one MI355X, term_updates[h]."""
    with np.load(path, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@dataclass
class SweepReport:
    n_query: int
    n_fallback: int
    fallback_h: int
    abs_err: np.ndarray        # per query point, vs the propagator
    frac_within_1e3: float
    max_err: float
    train_s: float
    infer_s: float
    fallback_s: float
    surrogate_s: float         # inference + fallback propagations
    plain_s: float             # propagator sweep over the same query points
    speedup: float
    train_loss: list
    abs_err_net: np.ndarray    # network alone (before fallback), per query point
    fallback_live: bool        # fallbacks propagated live (else charged at recorded times)
    tradeoff: list             # per disagreement bound: fallbacks, accuracy, speedup


def run_config5(grid: np.ndarray, h_values: np.ndarray, t_values: np.ndarray, plain_seconds: np.ndarray,
                n_train: int = 2048, seed: int = 31, members: int = 4, steps: int = 40000, lr: float = 1e-3,
                fallback_std: float = 1e-2, propagate=None, log=None) -> SweepReport:
    """Surrogate sweep of the (h, t) grid: train on ``n_train`` seed-sampled
    points, infer the rest, fall back to the propagator where the ensemble std
    exceeds ``fallback_std``.  ``grid[i, m]`` is the propagator value for h_i
    after m steps; ``plain_seconds[i]`` the propagator time for h_i over the
    full step range (the plain sweep's cost); ``propagate(h, steps)`` runs the
    propagator for fallbacks and returns the overlap series."""
    Hh, Tt = np.meshgrid(h_values, t_values, indexing="ij")
    steps_grid = np.rint(Tt / DT).astype(np.int64)
    Y = grid[np.arange(len(h_values))[:, None], steps_grid]
    flat = np.arange(Hh.size)
    rng = np.random.default_rng(seed)
    train_idx = np.sort(rng.choice(flat, size=n_train, replace=False))
    query_idx = np.setdiff1d(flat, train_idx)
    Xall = features(Hh.ravel(), Tt.ravel())
    Yall = Y.ravel()
    net = MLPEnsemble(Xall.shape[1], 256, members, max(len(train_idx), len(query_idx)), seed)
    t0 = time.perf_counter()
    loss = net.train(Xall[train_idx], Yall[train_idx].astype(np.float32), steps=steps, lr=lr)
    train_s = time.perf_counter() - t0
    if log:
        log(f"surrogate trained: {train_s:.2f}s, loss per member {loss.tolist()}")
    # timed surrogate sweep over the query points
    t0 = time.perf_counter()
    mean, sd = net.predict(Xall[query_idx])
    infer_s = time.perf_counter() - t0
    fb = sd > fallback_std
    pred = mean.astype(np.float64)
    err_net = np.abs(pred - Yall[query_idx])
    hi = (query_idx // len(t_values))
    fb_h = np.unique(hi[fb])
    t0 = time.perf_counter()
    for n_done, i in enumerate(fb_h):
        need = int(steps_grid.ravel()[query_idx[(hi == i) & fb]].max())
        ser = propagate(float(h_values[i]), need) if propagate else grid[i, :need + 1]
        if log and propagate:
            log(f"fallback {n_done + 1}/{len(fb_h)}: h={float(h_values[i]):.4f}, {need} steps, "
                f"{time.perf_counter() - t0:.1f}s")
        sel = (hi == i) & fb
        pred[sel] = ser[steps_grid.ravel()[query_idx[sel]]]
    fallback_s = time.perf_counter() - t0
    if propagate is None:
        # no live propagator: charge the recorded full-range time of each fallback h
        fallback_s = float(sum(plain_seconds[i] for i in fb_h))
    err = np.abs(pred - Yall[query_idx])
    q_h = np.unique(hi)
    plain_s = float(sum(plain_seconds[i] for i in q_h))
    # other disagreement bounds, fallbacks charged at the recorded propagation times
    tradeoff = []
    for bound in (1e-2, 5e-3, 2e-3, 1e-3, 5e-4):
        f = sd > bound
        hs = np.unique(hi[f])
        e = np.where(f, 0.0, err_net)  # a fallback point takes the propagator's value
        cost = infer_s + float(sum(plain_seconds[i] for i in hs))
        tradeoff.append({"std_bound": bound, "fallback_points": int(f.sum()), "fallback_h": int(len(hs)),
                         "frac_abs_err_le_1e-3": float(np.mean(e <= 1e-3)), "max_err": float(e.max()),
                         "speedup_recorded": plain_s / cost})
    surrogate_s = infer_s + fallback_s
    net.close()
    return SweepReport(len(query_idx), int(fb.sum()), len(fb_h), err, float(np.mean(err <= 1e-3)),
                       float(err.max()), train_s, infer_s, fallback_s, surrogate_s, plain_s,
                       plain_s / surrogate_s if surrogate_s > 0 else float("inf"), loss.tolist(),
                       err_net, propagate is not None, tradeoff)
