"""Batched propagation: many independent instances in one persistent kernel.

``Batch`` wraps the ``ppvm_batch_*`` C ABI.  ``run_hybrid`` is the scheduler
used for uneven batches: every instance starts in the persistent kernel with
an arena capped at ``batch_cap`` terms; instances that outgrow it are re-run
on the full-GPU engine (one handle per instance, several handles on their own
HIP streams driven from host threads), which is the better fit for large term
counts.  ``autotune`` picks the workgroup size, graph capture and the cap by
measuring candidate settings on the batch itself.
"""

from __future__ import annotations

import ctypes
import json
import os
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _native
from .ops import OP_DTYPE, pack_strings, words_for

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


class BatchInstance(ctypes.Structure):
    _fields_ = [
        ("n_qubits", ctypes.c_uint32),
        ("n_terms", ctypes.c_uint32),
        ("x", _u64p),
        ("z", _u64p),
        ("c", _f64p),
        ("ops", ctypes.c_void_p),
        ("n_ops", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("min_abs_coeff", ctypes.c_double),
        ("max_pauli_weight", ctypes.c_uint64),
        ("capacity", ctypes.c_uint64),
    ]


class BatchConfig(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_uint32), ("blocks", ctypes.c_uint32),
                ("segment_ops", ctypes.c_uint32), ("use_graph", ctypes.c_uint32)]


_declared = False


def _lib():
    global _declared
    L = _native.lib()
    if not _declared:
        L.ppvm_batch_last_error.restype = ctypes.c_char_p
        L.ppvm_batch_create.argtypes = [ctypes.POINTER(BatchInstance), ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.ppvm_batch_destroy.argtypes = [ctypes.c_void_p]
        L.ppvm_batch_set_capacity.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64]
        L.ppvm_batch_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchConfig), _u32p,
                                     ctypes.c_size_t, _f64p]
        L.ppvm_batch_results.argtypes = [ctypes.c_void_p, _u32p, _u64p, _f64p, _u64p, _u64p]
        L.ppvm_batch_latency.argtypes = [ctypes.c_void_p, _f64p]
        L.ppvm_batch_marks.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _f64p,
                                       ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.ppvm_batch_export.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _u64p, _f64p,
                                        ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.ppvm_batch_run1_noalloc.argtypes = [ctypes.POINTER(BatchInstance), ctypes.c_uint32, _u64p, _u64p,
                                              _f64p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _f64p,
                                              _u64p]
        _declared = True
    return L


def _check(rc: int):
    if rc != 0:
        msg = _lib().ppvm_batch_last_error().decode("utf-8", "replace")
        if rc == _native.PPVM_EINVAL:
            raise ValueError(msg)
        if rc == _native.PPVM_ENOMEM:
            raise MemoryError(msg)
        raise _native.PpvmError(rc, msg)


def marked_program(w) -> np.ndarray:
    """The workload's program with bit 0 of ``reserved`` set at its layer marks."""
    prog = np.array(w.program, dtype=OP_DTYPE, copy=True)
    if len(w.layer_marks):
        prog["reserved"][np.asarray(w.layer_marks, dtype=np.int64)] |= 1
    return prog


class Batch:
    """Owning wrapper of a ``ppvm_batch_handle`` built from workloads."""

    def __init__(self, workloads, capacity: int = 0):
        self.workloads = list(workloads)
        n = len(self.workloads)
        self._keep = []
        arr = (BatchInstance * n)()
        for i, w in enumerate(self.workloads):
            x, z = pack_strings([s for s, _ in w.observable], w.n_qubits)
            c = np.asarray([v for _, v in w.observable], dtype=np.float64)
            prog = marked_program(w)
            self._keep += [x, z, c, prog]
            inst = arr[i]
            inst.n_qubits = w.n_qubits
            inst.n_terms = len(c)
            inst.x = x.ctypes.data_as(_u64p)
            inst.z = z.ctypes.data_as(_u64p)
            inst.c = c.ctypes.data_as(_f64p)
            inst.ops = prog.ctypes.data
            inst.n_ops = len(prog)
            inst.min_abs_coeff = w.min_abs_coeff
            inst.max_pauli_weight = (1 << 64) - 1 if w.max_pauli_weight is None else w.max_pauli_weight
            inst.capacity = capacity
        h = ctypes.c_void_p()
        _check(_lib().ppvm_batch_create(arr, n, ctypes.byref(h)))
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None) is not None:
            _lib().ppvm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_capacity(self, i: int, cap: int):
        _check(_lib().ppvm_batch_set_capacity(self.h, i, int(cap)))

    def run(self, threads: int = 256, blocks: int = 0, segment_ops: int = 0, use_graph: bool = False,
            order=None) -> float:
        cfg = BatchConfig(threads, blocks, segment_ops, 1 if use_graph else 0)
        ms = ctypes.c_double()
        if order is None:
            _check(_lib().ppvm_batch_run(self.h, ctypes.byref(cfg), None, 0, ctypes.byref(ms)))
        else:
            o = np.ascontiguousarray(order, dtype=np.uint32)
            _check(_lib().ppvm_batch_run(self.h, ctypes.byref(cfg), o.ctypes.data_as(_u32p), len(o),
                                         ctypes.byref(ms)))
        return ms.value

    def results(self) -> dict:
        n = self.n
        st = np.zeros(n, dtype=np.uint32)
        nf = np.zeros(n, dtype=np.uint64)
        ov = np.zeros(n, dtype=np.float64)
        tu = np.zeros(n, dtype=np.uint64)
        pk = np.zeros(n, dtype=np.uint64)
        _check(_lib().ppvm_batch_results(self.h, st.ctypes.data_as(_u32p), nf.ctypes.data_as(_u64p),
                                         ov.ctypes.data_as(_f64p), tu.ctypes.data_as(_u64p),
                                         pk.ctypes.data_as(_u64p)))
        return {"status": st, "n_final": nf, "overlap": ov, "term_updates": tu, "peak": pk}

    def latency(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.float64)
        _check(_lib().ppvm_batch_latency(self.h, out.ctypes.data_as(_f64p)))
        return out

    def marks(self, i: int):
        cap = len(self.workloads[i].layer_marks)
        lens = np.zeros(cap, dtype=np.uint64)
        ovl = np.zeros(cap, dtype=np.float64)
        got = ctypes.c_size_t()
        _check(_lib().ppvm_batch_marks(self.h, i, lens.ctypes.data_as(_u64p), ovl.ctypes.data_as(_f64p),
                                       cap, ctypes.byref(got)))
        return lens[:got.value], ovl[:got.value]

    def export(self, i: int):
        w = self.workloads[i]
        W = words_for(w.n_qubits)
        n = int(self.results()["n_final"][i])
        x = np.zeros((n, W), dtype=np.uint64)
        z = np.zeros((n, W), dtype=np.uint64)
        c = np.zeros(n, dtype=np.float64)
        got = ctypes.c_size_t()
        _check(_lib().ppvm_batch_export(self.h, i, x.ctypes.data_as(_u64p), z.ctypes.data_as(_u64p),
                                        c.ctypes.data_as(_f64p), n, ctypes.byref(got)))
        return x, z, c


def run1_noalloc(w, capacity: int = 0, threads: int = 256, out_cap: int = 1 << 16):
    """One workload through the batched kernel with no device allocation
    (``ppvm_batch_run1_noalloc``).  Returns (x, z, c, overlap_last, term_updates)."""
    x, z = pack_strings([s for s, _ in w.observable], w.n_qubits)
    c = np.asarray([v for _, v in w.observable], dtype=np.float64)
    prog = marked_program(w)
    inst = BatchInstance()
    inst.n_qubits = w.n_qubits
    inst.n_terms = len(c)
    inst.x = x.ctypes.data_as(_u64p)
    inst.z = z.ctypes.data_as(_u64p)
    inst.c = c.ctypes.data_as(_f64p)
    inst.ops = prog.ctypes.data
    inst.n_ops = len(prog)
    inst.min_abs_coeff = w.min_abs_coeff
    inst.max_pauli_weight = (1 << 64) - 1 if w.max_pauli_weight is None else w.max_pauli_weight
    inst.capacity = capacity
    W = words_for(w.n_qubits)
    ox = np.zeros((out_cap, W), dtype=np.uint64)
    oz = np.zeros((out_cap, W), dtype=np.uint64)
    oc = np.zeros(out_cap, dtype=np.float64)
    n = ctypes.c_size_t()
    ovl = ctypes.c_double()
    tu = ctypes.c_uint64()
    _check(_lib().ppvm_batch_run1_noalloc(ctypes.byref(inst), threads, ox.ctypes.data_as(_u64p),
                                          oz.ctypes.data_as(_u64p), oc.ctypes.data_as(_f64p), out_cap,
                                          ctypes.byref(n), ctypes.byref(ovl), ctypes.byref(tu)))
    k = n.value
    return ox[:k], oz[:k], oc[:k], ovl.value, tu.value


@dataclass
class HybridResult:
    overlaps: list            # per instance: array of overlaps at the layer marks
    lens: list                # per instance: array of term counts at the layer marks
    term_updates: np.ndarray
    peak: np.ndarray
    on_engine: list           # indices re-run on the full-GPU engine
    batch_ms: float
    engine_s: float
    wall_s: float
    latency_s: np.ndarray     # per-instance completion time from the start of the run
    digests: list             # per engine instance: ppvm_fingerprint of its final terms (None in the kernel)


def _engine_run_digest(w):
    """_engine_run plus the ppvm_fingerprint digest of the final terms."""
    from .paulisum import PauliSum
    g = PauliSum.new(w.n_qubits, w.observable, min_abs_coeff=w.min_abs_coeff,
                     max_pauli_weight=w.max_pauli_weight)
    lens, ovls = [], []
    prev = 0
    for m in w.layer_marks:
        g.apply(w.program[prev:m + 1])
        prev = m + 1
        lens.append(len(g))
        ovls.append(g.overlap_with_zero())
    if prev < len(w.program):
        g.apply(w.program[prev:])
    g.synchronize()
    st = g.stats()
    return np.array(lens, dtype=np.uint64), np.array(ovls), st["term_updates"], st["peak_live"], g._h.fingerprint()


def _engine_run(w):
    """One workload on the full-GPU engine: term counts and overlaps at its
    layer marks, term updates, peak live terms."""
    lens, ovls, tu, pk, _ = _engine_run_digest(w)
    return lens, ovls, tu, pk


def run_hybrid(batch: Batch, threads: int = 256, use_graph: bool = True, engine_threads: int = 4,
               order=None, engine_set=None, segment_ops: int = 0, blocks: int = 0) -> HybridResult:
    """Run every instance.  Instances in ``engine_set`` start on the full-GPU
    engine at once (``engine_threads`` host threads, one stream each) while
    the persistent kernel runs the rest; instances that overflow the batch
    arena cap are re-run on the engine afterwards."""
    n = batch.n
    engine_set = sorted(set(engine_set or []), key=lambda i: -batch.workloads[i].n_gates)
    es = set(engine_set)
    if order is None:
        order = sorted(range(n), key=lambda i: -batch.workloads[i].n_gates)
    order = [i for i in order if i not in es]
    overlaps, lens = [None] * n, [None] * n
    tu = np.zeros(n, dtype=np.int64)
    pk = np.zeros(n, dtype=np.int64)
    lat = np.zeros(n)
    digests = [None] * n
    lock = threading.Lock()
    todo = list(engine_set)
    errors = []
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                if not todo:
                    return
                i = todo.pop(0)
            try:
                ln, ov, t, p, dg = _engine_run_digest(batch.workloads[i])
            except Exception as e:  # surfaced after join
                errors.append(e)
                return
            with lock:
                lens[i], overlaps[i], tu[i], pk[i], digests[i] = ln, ov, t, p, dg
                lat[i] = time.perf_counter() - t0

    ths = [threading.Thread(target=worker) for _ in range(max(1, min(engine_threads, len(todo))))]
    for t in ths:
        t.start()
    ms = batch.run(threads=threads, blocks=blocks, segment_ops=segment_ops, use_graph=use_graph,
                   order=order) if order else 0.0
    res = batch.results()
    dev_lat = batch.latency()
    redo = [i for i in order if res["status"][i] != 0]
    for i in order:
        if res["status"][i] == 0:
            lens[i], overlaps[i] = batch.marks(i)
            tu[i] = int(res["term_updates"][i])
            pk[i] = int(res["peak"][i])
            lat[i] = dev_lat[i]
    for t in ths:
        t.join()
    te = time.perf_counter()
    if redo:
        todo.extend(sorted(redo, key=lambda i: -batch.workloads[i].n_gates))
        ths = [threading.Thread(target=worker) for _ in range(max(1, min(engine_threads, len(redo))))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
    if errors:
        raise errors[0]
    t_end = time.perf_counter()
    return HybridResult(overlaps, lens, tu, pk, list(engine_set) + redo, ms, t_end - te, t_end - t0, lat, digests)


def autotune(workloads, caps=(1 << 16, 1 << 18, 1 << 20), threads=(256, 512, 1024), segments=(0, 256, 2048),
             blocks=(0, 1024), engine_threads=(8, 16), graphs=(True,), cache_path: str | None = None,
             log=None) -> dict:
    """Choose the batched run's settings by measuring them on the batch itself.

    Searched: the per-instance arena cap (instances that outgrow it run on the
    full-GPU engine, concurrently with the persistent kernel), the workgroup
    size (the tile of terms one pass of a workgroup covers), the segment size
    (ops per kernel launch of the captured launch sequence; 0 = whole
    programs), the persistent grid (0 = two workgroups per CU) and the engine's
    host threads.  For each cap a first run through the kernel alone learns
    which instances outgrow it; then one setting is varied at a time from the
    current best (coordinate search), so the trials grow with the sum, not the
    product, of the candidate counts.  The winner (with its engine set) is
    cached in ``cache_path`` (JSON) keyed by a signature of the workloads."""
    sig = f"{len(workloads)}:{sum(w.n_gates for w in workloads)}:{sum(w.n_qubits for w in workloads)}"
    if cache_path and os.path.exists(cache_path):
        try:
            cached = json.load(open(cache_path))
            if sig in cached:
                return cached[sig]
        except (OSError, ValueError):
            pass
    trials = []
    best = None

    def trial(b, big, cfg):
        nonlocal best
        r = run_hybrid(b, threads=cfg["threads"], use_graph=cfg["graph"], engine_threads=cfg["engine_threads"],
                       engine_set=big, segment_ops=cfg["segment_ops"], blocks=cfg["blocks"])
        t = dict(cfg, wall_s=r.wall_s, batch_ms=r.batch_ms, on_engine=len(r.on_engine), engine_set=big)
        trials.append(t)
        if log:
            log("autotune " + " ".join(f"{k}={cfg[k]}" for k in cfg) +
                f": {r.wall_s:.3f}s (kernel {r.batch_ms:.0f} ms, {len(r.on_engine)} on engine)")
        if best is None or r.wall_s < best["wall_s"]:
            best = dict(t)
        return r.wall_s

    base = {"threads": threads[-1], "graph": graphs[0], "engine_threads": engine_threads[-1],
            "segment_ops": segments[0], "blocks": blocks[0]}
    for cap in caps:
        b = Batch(workloads, capacity=cap)
        first = run_hybrid(b, threads=base["threads"], use_graph=False)
        big = sorted(first.on_engine)
        cur = dict(base, cap=cap)
        cur_s = trial(b, big, cur)
        for key, cands in (("threads", threads), ("segment_ops", segments), ("blocks", blocks),
                           ("engine_threads", engine_threads), ("graph", graphs)):
            for v in cands:
                if v == cur[key]:
                    continue
                cfg = dict(cur, **{key: v})
                s_ = trial(b, big, cfg)
                if s_ < cur_s:
                    cur, cur_s = cfg, s_
        b.close()
    out = dict(best, trials=[{k: v for k, v in t.items() if k != "engine_set"} for t in trials],
               signature=sig)
    if cache_path:
        try:
            cached = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        except (OSError, ValueError):
            cached = {}
        cached[sig] = out
        with open(cache_path, "w") as f:
            json.dump(cached, f, indent=1)
    return out
