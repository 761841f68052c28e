"""Learned surrogate for parameter sweeps of the lattice TFIM.

An ensemble of MLPs (inputs: h, t, steps and sin/cos features; hidden layers
256-256-256 with GELU; scalar output; BF16 weights and activations with FP32
accumulation on MFMA) is trained on the GPU, by the library's own forward and
backward kernels, on expectation values produced by the GPU propagator.  At
query time the ensemble mean is used where the ensemble agrees (standard
deviation within a bound calibrated on held-out validation points);
elsewhere the propagator is run (one propagation per distinct h, prefix
reuse over t).
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
class SweepSplit:
    """Grid points of a sweep split for training, calibration and query.

    Points are flat indices into the (h, t) grid.  ``pair`` names the physical
    point (h index, step count): grid columns with equal ``round(t / dt)`` are
    the same point.  Training points are the seed-sampled ones; a
    validation split of their distinct pairs is held out of fitting and used
    only to calibrate the fallback bound.  A query point is *known* when its
    pair is a training pair (its value is then the training target) and *new*
    otherwise; new pairs share nothing with the fitted or validation pairs."""
    steps: np.ndarray      # steps per flat grid point
    hidx: np.ndarray       # h index per flat grid point
    pair: np.ndarray       # pair id per flat grid point
    train: np.ndarray      # seed-sampled training points
    fit: np.ndarray        # training points the ensemble is fitted on
    val: np.ndarray        # training points held out for calibration (pairs disjoint from fit)
    query: np.ndarray      # all other grid points
    known: np.ndarray      # per query point: pair is a training pair


def split_sweep(h_values, t_values, n_train: int = 2048, seed: int = 31, val_frac: float = 0.2,
                dt: float = DT) -> SweepSplit:
    nh, nt = len(h_values), len(t_values)
    steps = np.broadcast_to(np.rint(np.asarray(t_values) / dt).astype(np.int64), (nh, nt)).ravel()
    hidx = np.repeat(np.arange(nh), nt)
    pair = hidx * (int(steps.max()) + 1) + steps
    flat = np.arange(nh * nt)
    train = np.sort(np.random.default_rng(seed).choice(flat, size=n_train, replace=False))
    query = np.setdiff1d(flat, train)
    tpairs = np.unique(pair[train])
    vpairs = np.random.default_rng(seed + 1).choice(tpairs, size=int(round(val_frac * len(tpairs))),
                                                    replace=False)
    in_val = np.isin(pair[train], vpairs)
    return SweepSplit(steps, hidx, pair, train, train[~in_val], train[in_val], query,
                      np.isin(pair[query], tpairs))


def calibrate_bound(sd: np.ndarray, err: np.ndarray, bar: float) -> float:
    """Largest ensemble-std bound under which every validation point is within
    ``bar`` of the propagator: the std of the last point, in increasing std,
    before the first one outside the bar.  -1 when even the most confident
    point misses (every query point then falls back)."""
    order = np.argsort(sd, kind="stable")
    bad = np.nonzero(err[order] > bar)[0]
    k = len(order) if len(bad) == 0 else int(bad[0])
    return float(sd[order[k - 1]]) if k > 0 else -1.0


@dataclass
class SweepReport:
    n_query: int
    n_known: int               # query points whose (h, steps) is a training pair: answered by lookup
    n_new: int
    bound: float               # fallback bound on the ensemble std (calibrated unless given)
    val_points: int
    val_accepted: int          # validation points under the bound (all within the bar)
    n_accepted: int            # new query points answered by the network
    n_fallback: int            # new query points answered by a fallback propagation
    fallback_h: int
    abs_err: np.ndarray        # per query point, vs this run's propagator values
    frac_within_bar: float
    max_err: float
    train_s: float
    infer_s: float
    fallback_s: float
    surrogate_s: float         # whole query phase: lookup, inference, fallback propagations
    plain_s: float             # live plain sweep of the same query points
    speedup: float
    train_loss: list
    abs_err_net: np.ndarray    # network alone on the new query points
    grid_max_diff: float       # |live plain sweep - training grid| over the points both hold
    tradeoff: list             # other bounds, fallbacks charged at this run's plain prefix times


def run_sweep(grid: np.ndarray, h_values, t_values, propagate, n_train: int = 2048, seed: int = 31,
              members: int = 4, steps: int = 12000, lr: float = 2e-3, bar: float = 1e-3,
              val_frac: float = 0.2, bound: float | None = None, log=None) -> SweepReport:
    """Surrogate sweep of an (h, t) grid against the plain propagator sweep,
    both timed live in this call.

    ``grid[i, m]`` holds the propagator's value for ``h_values[i]`` after m
    steps; only the training points are read from it.  ``propagate(h, m)``
    runs the propagator and returns (values after 0..m steps, cumulative
    seconds after 0..m steps).

    1. Plain sweep (timed): one propagation per h value that holds a query
       point, to its largest query step.  Its values judge every error below.
    2. Fit the ensemble on the training points outside the validation split.
    3. Bound (unless ``bound`` is given): the largest ensemble std under which
       every validation point is within ``bar``.
    4. Query (timed): known pairs by lookup; new pairs by the ensemble where
       its std is within the bound; for every h holding a new pair above the
       bound, one propagation to the largest such step, whose values answer
       every query point of that h it covers."""
    h_values = np.asarray(h_values, dtype=np.float64)
    sp = split_sweep(h_values, t_values, n_train, seed, val_frac)
    nq = len(sp.query)
    qh, qs = sp.hidx[sp.query], sp.steps[sp.query]
    # 1. plain sweep
    q_h = np.unique(qh)
    need = {int(i): int(qs[qh == i].max()) for i in q_h}
    truth, cum = {}, {}
    t0 = time.perf_counter()
    for n_done, i in enumerate(q_h):
        truth[int(i)], cum[int(i)] = propagate(float(h_values[i]), need[int(i)])
        if log and (n_done % 8 == 7 or n_done == len(q_h) - 1):
            log(f"plain sweep {n_done + 1}/{len(q_h)} h values, {time.perf_counter() - t0:.1f}s")
    plain_s = time.perf_counter() - t0
    ref = np.array([truth[int(i)][m] for i, m in zip(qh, qs)])
    grid_diff = max(float(np.max(np.abs(truth[i] - grid[i, :len(truth[i])]))) for i in truth)
    # 2. fit
    Xall = features(h_values[sp.hidx], sp.steps * DT)
    Yall = grid[sp.hidx, sp.steps]
    net = MLPEnsemble(Xall.shape[1], 256, members, max(len(sp.fit), nq), seed)
    t0 = time.perf_counter()
    loss = net.train(Xall[sp.fit], Yall[sp.fit].astype(np.float32), steps=steps, lr=lr)
    train_s = time.perf_counter() - t0
    if log:
        log(f"ensemble fitted on {len(sp.fit)} points: {train_s:.2f}s, loss per member {loss.tolist()}")
    # 3. bound
    vm, vsd = net.predict(Xall[sp.val])
    verr = np.abs(vm.astype(np.float64) - Yall[sp.val])
    if bound is None:
        bound = calibrate_bound(vsd, verr, bar)
    val_acc = int(np.sum(vsd <= bound))
    # 4. query
    new = ~sp.known
    lut = dict(zip(sp.pair[sp.train].tolist(), Yall[sp.train].tolist()))
    pred = np.zeros(nq)
    t0 = time.perf_counter()
    pred[sp.known] = [lut[p] for p in sp.pair[sp.query[sp.known]].tolist()]
    mean, sd = net.predict(Xall[sp.query[new]])
    infer_s = time.perf_counter() - t0
    pred[new] = mean
    fb = np.zeros(nq, dtype=bool)
    fb[np.nonzero(new)[0][sd > bound]] = True
    fb_h = np.unique(qh[fb])
    t1 = time.perf_counter()
    for n_done, i in enumerate(fb_h):
        top = int(qs[fb & (qh == i)].max())
        ser, _ = propagate(float(h_values[i]), top)
        cover = (qh == i) & (qs <= top) & ~sp.known
        pred[cover] = ser[qs[cover]]
        if log:
            log(f"fallback {n_done + 1}/{len(fb_h)}: h={h_values[i]:.4f} to {top} steps, "
                f"{time.perf_counter() - t1:.1f}s")
    fallback_s = time.perf_counter() - t1
    surrogate_s = time.perf_counter() - t0
    net.close()
    err = np.abs(pred - ref)
    err_net = np.abs(mean.astype(np.float64) - ref[new])
    # other bounds, each fallback h charged at this run's plain-sweep time to its largest fallback step
    tradeoff = []
    newq = np.nonzero(new)[0]
    for b in sorted({bound, 1e-2, 5e-3, 2e-3, 1e-3, 5e-4, 0.0}, reverse=True):
        f = np.zeros(nq, dtype=bool)
        f[newq[sd > b]] = True
        e = err.copy()
        e[newq] = err_net
        cost = infer_s
        for i in np.unique(qh[f]):
            top = int(qs[f & (qh == i)].max())
            e[(qh == i) & (qs <= top) & ~sp.known] = 0.0
            cost += float(cum[int(i)][top])
        tradeoff.append({"std_bound": b, "fallback_points": int(f.sum()), "fallback_h": int(len(np.unique(qh[f]))),
                         "frac_abs_err_le_bar": float(np.mean(e <= bar)), "max_err": float(e.max()),
                         "speedup": plain_s / cost})
    return SweepReport(nq, int(sp.known.sum()), int(new.sum()), float(bound), len(sp.val), val_acc,
                       int(new.sum() - fb.sum()), int(fb.sum()), len(fb_h), err, float(np.mean(err <= bar)),
                       float(err.max()), train_s, infer_s, fallback_s, surrogate_s, plain_s,
                       plain_s / surrogate_s if surrogate_s > 0 else float("inf"), loss.tolist(), err_net,
                       grid_diff, tradeoff)
