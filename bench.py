"""Headline benchmark: Pauli-term updates/s of the 10x10 lattice TFIM on one MI355X.

One "step" is one full backward propagation of the lattice workload (10x10
open lattice, h=3.04, J=1, dt=0.05, 20 Trotter steps = 5,600 rotation gates,
threshold 1e-5, max weight 10, truncation after every gate) from the
observable Z_(5,5).  Inputs are resident in HBM before the timed region: each
timed step runs on a handle created (and sized for the peak term count) before
the clock starts.  A term update is one live input term processed by one gate
(U = sum over gates of live input terms; identical on CPU and GPU because the
term sets are identical).

Prints one JSON line.  Extra objects:
  roofline     dominant kernel's algorithmic bytes / its average launch time
               (HIP events on the handle's stream, profiled pass);
  cpu_baseline the CPU oracle on a bounded prefix of the same workload, timed
               on this host in the same run: the better of the single-thread
               oracle and its sharded multi-thread variant (cores stated);
  parity       the GPU term set on that prefix vs the oracle's (bit-identical).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Pauli-term updates/s, 10x10 TFIM vs CPU; surrogate sweep speedup @ |err|<=1e-3"


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _make(w, cap: int = 0):
    from ppvm_amd import PauliSum
    return PauliSum.new(w.n_qubits, w.observable, min_abs_coeff=w.min_abs_coeff,
                        max_pauli_weight=w.max_pauli_weight, capacity_hint=cap)


# Timed phases of the engine (HIP events per launch); "pairing" is k_rot_pair
# on small gates and k_pair_hist + k_pair_colscan + k_pair_scatter_tiled +
# k_pair_lds on large ones.
KERNELS = ["k_rot_count", "k_rot_gather", "pairing", "k_emit_count", "k_emit_write", "k_map",
           "k_finish", "other"]
PAIRING_KERNELS = ("k_rot_pair", "k_pair_hist", "k_pair_colscan", "k_pair_scatter_tiled", "k_pair_lds")


def phase_bytes(cnt: dict, W: int, nq: int) -> dict:
    """Algorithmic HBM bytes per kernel, summed over the run (S = 16W + 8 B per
    term: key words and coefficient; R = 16 B per worklist record: coefficient,
    arena index, key facts; N = scanned entries, N_a = anticommuting terms,
    N_p = merged pairs, N_c = appended children, N_t = entries transposed into
    the bit-sliced planes):

    k_rot_count  : 1 bit per bit-sliced plane the generator needs for every
                   entry + N/8 B of ballot masks + N_t (16W + 2 nq / 8) for the
                   entries tested from their keys and transposed.
    k_rot_gather : N/8 B masks + N_a (S + R + 8): key and coefficient read,
                   record and pair-key hash written.
    pairing      : N_a 8 B hashes read + N_p 2 (R + 16W + 16): both members'
                   records and arena keys read, both records rewritten (the
                   minimum of either pairing implementation; the partition's
                   own traffic shows in the measured bytes).
    k_emit_count : N_a R records read + N_a 8 B coefficients written (scaled
                   parents and merged pairs) + N_a/8 masks.
    k_emit_write : N_a/8 masks + N_c (R + 16W + S): record and parent key read,
                   child term written.
    """
    S = 16 * W + 8
    R = 16
    N = cnt["scanned"]
    na = cnt["wl_total"]
    npair = cnt["pairs_total"]
    nch = cnt["children_total"]
    return {
        "k_rot_count": cnt["scan_planes"] / 8 + N / 8 + cnt["transposed"] * (16 * W + 2 * nq / 8),
        "k_rot_gather": N / 8 + na * (S + R + 8),
        "pairing": na * 8 + npair * 2 * (R + 16 * W + 16),
        "k_emit_count": na * R + na * 8 + na / 8,
        "k_emit_write": na / 8 + nch * (R + 16 * W + S),
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_lattice_tfim_10x10.txt")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel), or None.
    For the "pairing" phase: the bytes of all its kernels per gate."""
    rows = {}
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                # the first table averages over every dispatch of the run; later
                # tables (a subset of gates) are not per-launch averages of it
                if rows and line.startswith("# HBM traffic"):
                    break
                parts = line.split()
                if len(parts) == 5 and parts[0].startswith("k_"):
                    rows[parts[0]] = (int(parts[1]), float(parts[4]))
    except (OSError, ValueError):
        return None
    if kernel == "pairing":
        gates = rows.get("k_rot_gather", (0, 0.0))[0]
        if not gates:
            return None
        return sum(n * b for k, (n, b) in rows.items() if k in PAIRING_KERNELS) / gates
    return rows[kernel][1] if kernel in rows else None


# DevState counters that accumulate over a handle's life (the rest are levels)
CUMULATIVE = ("term_updates", "gates", "rotations", "scanned", "wl_total", "pairs_total", "children_total",
              "kills_total", "map_entries", "map_launches", "scan_planes", "transposed", "pair_fallback")


def counter_delta(after: dict, before: dict) -> dict:
    return {k: (after[k] - before[k]) if k in CUMULATIVE else after[k] for k in after}


def run(args) -> dict:
    from ppvm_amd import _native
    from ppvm_amd import workloads as WL

    info = _native.device_info()
    _log(f"device: {info}")
    w = WL.lattice_tfim()
    W = 2
    # warm-up: the first propagation grows the handle's arenas to the peak
    # size and loads the kernels; every later propagation reuses the same
    # arenas (reset = clear + upload of the one-term observable)
    t0 = time.perf_counter()
    g = _make(w)
    g.apply(w.program)
    g.synchronize()
    warm_s = time.perf_counter() - t0
    c0 = g._h.counters()
    U = c0["term_updates"]
    peak = c0["peak_live"]
    cap = int(c0["capacity"])
    ovl = g.overlap_with_zero()
    n_final = len(g)
    _log(f"warm-up run: {warm_s:.2f}s U={U} peak_live={peak} final={n_final} cap={cap} overlap={ovl:.12f}")
    for _ in range(max(0, args.warmup - 1)):
        g.reset()
        g.apply(w.program)
        g.synchronize()

    # timed region: K full propagations on the pre-sized handle; each step's
    # clock runs from a synchronised, freshly reset handle (observable already
    # in HBM) to the synchronised end of the propagation
    elapsed = 0.0
    step_s = []
    for _ in range(args.steps):
        g.reset()
        before = g._h.counters()["term_updates"]  # synchronises
        t0 = time.perf_counter()
        g.apply(w.program)
        g.synchronize()
        dt = time.perf_counter() - t0
        elapsed += dt
        step_s.append(dt)
        assert g._h.counters()["term_updates"] - before == U, "non-deterministic term-update count"
        assert abs(g.overlap_with_zero() - ovl) <= 1e-10
    value = U * args.steps / elapsed
    _log(f"timed: {args.steps} x {elapsed / args.steps:.3f}s  -> {value:.4e} term-updates/s")

    # profiled pass: per-kernel HIP-event timing on the handle's stream
    g.reset()
    cb = g._h.counters()
    g._h.profile(True)
    g.apply(w.program)
    g.synchronize()
    ms, launches = g._h.profile_read()
    g._h.profile(False)
    cnt = counter_delta(g._h.counters(), cb)
    pb = phase_bytes(cnt, W, w.n_qubits)
    kern = {}
    for i, nm in enumerate(KERNELS):
        if launches[i]:
            kern[nm] = {"ms_total": float(ms[i]), "launches": int(launches[i]),
                        "avg_us": float(ms[i]) * 1e3 / int(launches[i])}
            if nm in pb:
                kern[nm]["alg_bytes_total"] = float(pb[nm])
                kern[nm]["achieved_GBs"] = pb[nm] / (kern[nm]["ms_total"] * 1e-3) / 1e9
                tr = measured_traffic(nm)
                if tr is not None:
                    # rocprofv3 FETCH_SIZE + WRITE_SIZE per launch (committed profile of
                    # this workload) over this run's time per launch
                    kern[nm]["measured_hbm_GBs"] = tr * int(launches[i]) / (kern[nm]["ms_total"] * 1e-3) / 1e9
    # the dominant kernel: the single kernel with the largest total time (the
    # pairing phase is several kernels, each well below the gather)
    dom = max((k for k in kern if k.startswith("k_") and k in pb), key=lambda k: kern[k]["ms_total"])
    _log("kernels: " + json.dumps(kern))
    dk = kern[dom]
    achieved = dk.get("achieved_GBs", 0.0)
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(dom),
        "traffic_source": "profiles/hbm_traffic_lattice_tfim_10x10.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE "
                          "per dispatch, calibrated)",
        "avg_launch_us": round(dk["avg_us"], 3),
        "alg_bytes_per_launch": round(dk.get("alg_bytes_total", 0.0) / dk["launches"], 1),
    }

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "term-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "step_ms": {"min": min(step_s) * 1e3, "max": max(step_s) * 1e3},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "lattice_tfim: 10x10 open TFIM, J=1, h=3.04, dt=0.05, 20 Trotter steps, "
                        "Z(5,5), threshold 1e-5, max weight 10, truncate after every gate",
            "gates": int(w.n_gates),
            "term_updates_per_step": int(U),
            "peak_live_terms": int(peak),
            "final_terms": int(n_final),
            "overlap_with_zero": ovl,
            "device": info,
            "arena_capacity": cap,
        },
        "roofline": roofline,
        "kernels": kern,
    }

    # CPU baseline + parity on a bounded prefix of the same workload
    if not args.no_cpu:
        from oracle.pyoracle import OracleSum
        ps = args.cpu_prefix_steps
        p = WL.truncate_steps(w, ps)
        ref = OracleSum(p.n_qubits, p.min_abs_coeff, p.max_pauli_weight)
        ref.add_strings([s for s, _ in p.observable], [c for _, c in p.observable])
        t0 = time.perf_counter()
        ref.apply(p.program)
        cpu_s = time.perf_counter() - t0
        cpu_rate = ref.term_updates() / cpu_s
        # the GPU on the same prefix, same pre-sized handle (best of 3 runs)
        gpu_s = None
        for _ in range(3):
            g.reset()
            cb = g._h.counters()["term_updates"]
            t0 = time.perf_counter()
            g.apply(p.program)
            g.synchronize()
            dt = time.perf_counter() - t0
            gpu_s = dt if gpu_s is None else min(gpu_s, dt)
            gu = g._h.counters()["term_updates"] - cb
        gp = g
        gx, gz, gc = gp._export()
        ox, oz, oc = ref.export()
        same = (len(gc) == len(oc) and np.array_equal(gx, ox) and np.array_equal(gz, oz)
                and np.array_equal(gc.view(np.uint64), oc.view(np.uint64)))
        # (ii) sharded multi-thread variant on the host's cores (same prefix)
        threads = max(1, min(16, os.cpu_count() or 1))
        shd = OracleSum(p.n_qubits, p.min_abs_coeff, p.max_pauli_weight)
        shd.add_strings([s for s, _ in p.observable], [c for _, c in p.observable])
        t0 = time.perf_counter()
        shd.apply_sharded(p.program, threads)
        shd_s = time.perf_counter() - t0
        shd_rate = shd.term_updates() / shd_s
        shd_same = dict(shd.terms()) == dict(ref.terms())
        single = {"value": cpu_rate, "cores": 1, "seconds": cpu_s}
        sharded = {"value": shd_rate, "cores": threads, "seconds": shd_s, "same_terms": bool(shd_same)}
        best = sharded if shd_rate > cpu_rate else single
        result["cpu_baseline"] = {
            "value": best["value"],
            "unit": "term-updates/s",
            "cores": best["cores"],
            "kind": "port",
            "sample": f"first {ps} of 20 Trotter steps of the same lattice workload "
                      f"({len(p.program)} ops, U={ref.term_updates()}); better of the single-thread "
                      f"oracle ({cpu_s:.1f}s) and its {threads}-thread sharded variant ({shd_s:.1f}s)",
            "single_thread": single,
            "sharded": sharded,
            "host_cpus": os.cpu_count(),
            "cpu_model": cpu_model(),
            "gpu_same_prefix": gu / gpu_s,
            "speedup_same_prefix": (gu / gpu_s) / best["value"],
        }
        cpu_rate = best["value"]
        result["parity"] = {
            "prefix_steps": ps,
            "terms": int(len(oc)),
            "bit_identical": bool(same),
            "term_updates_equal": bool(gu == ref.term_updates()),
            "overlap_abs_diff": abs(gp.overlap_with_zero() - ref.overlap_with_zero()),
        }
        # like with like: GPU and CPU rates over the same prefix of the workload
        result["speedup_vs_cpu"] = (gu / gpu_s) / cpu_rate
        result["speedup_vs_cpu_basis"] = (f"GPU and better CPU baseline on the same {ps}-step prefix; "
                                          "the full-run GPU rate (value) over the prefix CPU rate is "
                                          f"{value / cpu_rate:.1f}x")
        _log(f"cpu: single {single['value']:.4e}/s, sharded x{threads} {shd_rate:.4e}/s over {ps} steps; "
             f"parity={same}")
    return result


def run_batched(args) -> dict:
    """256 independent 1D TFIM chains (L~U{32..128}, h~U[0.5,1.5], steps~U{5..30},
    seed 21, threshold 1e-5): autotuned hybrid of the persistent batched kernel
    and the full-GPU engine.  Reports aggregate term-updates/s, p50/p99
    per-instance latency, per-instance overlap agreement with sequential plain
    engine runs (sample) and with the CPU oracle (sample), and the CPU oracle
    rate on a bounded sample of instances."""
    from ppvm_amd import workloads as WL
    from ppvm_amd.batch import Batch, autotune, run_hybrid

    ws = WL.batched_chains()
    t0 = time.perf_counter()
    best = autotune(ws, log=_log)
    tune_s = time.perf_counter() - t0
    b = Batch(ws, capacity=best["cap"])
    runs = []
    for _ in range(max(1, args.steps)):
        r = run_hybrid(b, threads=best["threads"], use_graph=best["graph"],
                       engine_threads=best["engine_threads"], engine_set=best["engine_set"],
                       segment_ops=best["segment_ops"], blocks=best["blocks"])
        runs.append(r)
        _log(f"batched run: {r.wall_s:.3f}s kernel {r.batch_ms:.0f} ms, {len(r.on_engine)} on engine")
    r = min(runs, key=lambda x: x.wall_s)
    U = int(r.term_updates.sum())
    # agreement with sequential plain engine runs and with the oracle, on a sample
    from ppvm_amd.batch import _engine_run
    rng = np.random.default_rng(0)
    sample = sorted(rng.choice(len(ws), size=min(args.verify, len(ws)), replace=False).tolist())
    max_dov, len_ok = 0.0, True
    for i in sample:
        ln, ov, _, _ = _engine_run(ws[i])
        len_ok &= bool(np.array_equal(np.asarray(r.lens[i], dtype=np.uint64), ln))
        max_dov = max(max_dov, float(np.max(np.abs(np.asarray(r.overlaps[i]) - ov))))
    # every instance against the oracle's committed per-layer marks
    fix_path = os.path.join(ROOT, "tests", "golden", "batched_chains_marks.json")
    fixture = None
    if os.path.exists(fix_path):
        with open(fix_path) as f:
            fixture = json.load(f)["instances"]
    if fixture is not None:
        lens_all = all(np.array_equal(np.asarray(r.lens[i], dtype=np.uint64),
                                      np.asarray(fixture[i]["lens"], dtype=np.uint64)) for i in range(len(ws)))
        ovl_all = max(float(np.max(np.abs(np.asarray(r.overlaps[i]) - np.asarray(fixture[i]["overlaps"]))))
                      for i in range(len(ws)))
        tu_all = all(int(r.term_updates[i]) == fixture[i]["term_updates"] for i in range(len(ws)))
    in_kernel = [i for i in range(len(ws)) if i not in set(r.on_engine)]
    out = {
        "metric": "batched term-updates/s, 256 TFIM chains",
        "value": U / r.wall_s,
        "unit": "term-updates/s",
        "n_gpus": 1,
        "steps": len(runs),
        "ms_per_step": r.wall_s * 1e3,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "256 1D TFIM chains, L~U{32..128}, h~U[0.5,1.5], J=1, dt=0.1, "
                               "steps~U{5..30} (seed 21), Z at L//2, threshold 1e-5",
                   "gates": int(sum(w.n_gates for w in ws)), "term_updates": U,
                   "autotuned": {k: v for k, v in best.items() if k not in ("trials", "engine_set")},
                   "autotune_s": tune_s, "autotune_trials": best.get("trials")},
        "latency_s": {"p50": float(np.percentile(r.latency_s, 50)),
                      "p99": float(np.percentile(r.latency_s, 99)),
                      "max": float(np.max(r.latency_s)),
                      "persistent_kernel_instances": len(in_kernel),
                      "persistent_kernel_p50": float(np.percentile(r.latency_s[in_kernel], 50)) if in_kernel else None,
                      "persistent_kernel_p99": float(np.percentile(r.latency_s[in_kernel], 99)) if in_kernel else None},
        "kernel_ms": r.batch_ms,
        "on_engine": len(r.on_engine),
        "parity_vs_sequential_engine": {"instances": len(sample), "layer_counts_equal": len_ok,
                                        "max_overlap_abs_diff": max_dov},
    }
    if fixture is not None:
        out["parity_vs_oracle_marks"] = {"instances": len(ws), "layer_counts_equal": bool(lens_all),
                                         "term_updates_equal": bool(tu_all), "max_overlap_abs_diff": ovl_all}
    if not args.no_cpu:
        from oracle.pyoracle import OracleSum
        # random-order sample of batch instances within a CPU time budget; an
        # instance is taken if its term updates (known from the GPU run) fit the
        # remaining budget at the rate measured so far
        rng = np.random.default_rng(1)
        budget, rate = args.cpu_budget, None
        tu, t_cpu, max_dor, taken = 0, 0.0, 0.0, []
        for i in rng.permutation(len(ws)).tolist():
            if i in r.on_engine:
                continue
            ui = int(r.term_updates[i])
            if rate is not None and ui / rate > budget - t_cpu:
                continue
            w = ws[i]
            o = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
            o.add_strings([s for s, _ in w.observable], [c for _, c in w.observable])
            t1 = time.perf_counter()
            lens, ovl = o.run_marked(w.program, w.layer_marks)
            t_cpu += time.perf_counter() - t1
            tu += o.term_updates()
            rate = tu / t_cpu
            taken.append(i)
            max_dor = max(max_dor, float(np.max(np.abs(np.asarray(r.overlaps[i]) - ovl))))
            if t_cpu >= budget or len(taken) >= args.cpu_instances:
                break
        # the same instances spread over the host's cores (one oracle per thread;
        # the oracle's C calls run outside the interpreter lock)
        from concurrent.futures import ThreadPoolExecutor
        threads = max(1, min(16, os.cpu_count() or 1))

        def one(i):
            w = ws[i]
            o = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
            o.add_strings([s for s, _ in w.observable], [c for _, c in w.observable])
            o.run_marked(w.program, w.layer_marks)
            return o.term_updates()
        t1 = time.perf_counter()
        with ThreadPoolExecutor(threads) as pool:
            tu_par = sum(pool.map(one, sorted(taken, key=lambda i: -int(r.term_updates[i]))))
        par_s = time.perf_counter() - t1
        single = {"value": tu / t_cpu, "cores": 1, "seconds": t_cpu}
        parallel = {"value": tu_par / par_s, "cores": threads, "seconds": par_s}
        best = parallel if parallel["value"] > single["value"] else single
        out["cpu_baseline"] = {"value": best["value"], "unit": "term-updates/s", "cores": best["cores"],
                               "kind": "port",
                               "sample": f"{len(taken)} batch instances in random order (seed 1) within a "
                                         f"{budget:.0f}s single-thread budget: U={tu} ({tu / U:.2%} of the "
                                         f"batch); better of one thread ({t_cpu:.1f}s) and {threads} threads, "
                                         f"one instance each ({par_s:.1f}s)",
                               "single_thread": single, "parallel": parallel, "host_cpus": os.cpu_count(),
                               "cpu_model": cpu_model()}
        out["parity_vs_oracle"] = {"instances": len(taken), "max_overlap_abs_diff": max_dor}
    return out


def run_surrogate(args) -> dict:
    """Learned-surrogate sweep of the lattice TFIM over 64 h in [0.5, 4] x 64 t in
    [0, 2] (dt 0.05, steps = round(t/dt)), everything timed live in this run.
    The ensemble (4 MLPs, 256-256-256 GELU, BF16 MFMA / FP32 accumulation) is
    fitted on the GPU on the 2048 seed-31 training points of the shipped
    propagator grid (less a validation split of 20% of their (h, steps) pairs,
    held out to calibrate the fallback bound on the ensemble std so that every
    validation point under it is within 1e-3).  The other 2048 points are the
    query: known (h, steps) pairs by lookup, new pairs by the ensemble under
    the bound, the propagator for every h holding a new pair above it.  The
    comparator is the plain sweep of the same query points (one prefix-reuse
    propagation per h), timed in the same run; its values judge the errors and
    must reproduce the shipped grid."""
    from ppvm_amd import sweep as SW
    from ppvm_amd.surrogate import load_grid, run_sweep

    d = load_grid()

    def propagate(h, m):
        ser, _, _, secs = SW.lattice_series(h, m)
        return ser, secs

    rep = run_sweep(d["grid"], d["h"], d["t"], propagate, bound=args.surrogate_bound, log=_log)
    err = rep.abs_err
    return {
        "metric": "surrogate sweep speedup vs plain GPU sweep @ |err|<=1e-3",
        "value": rep.speedup,
        "unit": "x",
        "n_gpus": 1,
        "steps": 1,
        "ms_per_step": rep.surrogate_s * 1e3,
        "higher_is_better": True,
        "dtype": "bf16 (f32 accumulation); propagator f64",
        "data": "synthetic",
        "config": {"workload": "lattice_tfim sweep: 64 h in [0.5,4] x 64 t in [0,2], dt 0.05; "
                               "4-member ensemble 256-256-256 GELU, 2048 training points (seed 31), "
                               "2048 query points"},
        "query_points": rep.n_query,
        "known_pairs": rep.n_known,
        "new_pairs": rep.n_new,
        "std_bound": rep.bound,
        "bound_calibrated": args.surrogate_bound is None,
        "validation": {"points": rep.val_points, "under_bound": rep.val_accepted},
        "network_answers": rep.n_accepted,
        "fallback_points": rep.n_fallback,
        "fallback_h_values": rep.fallback_h,
        "frac_abs_err_le_1e-3": rep.frac_within_bar,
        "abs_err": {"median": float(np.median(err)), "p90": float(np.percentile(err, 90)),
                    "p99": float(np.percentile(err, 99)), "max": rep.max_err},
        "network_only_new_pairs": {"frac_abs_err_le_1e-3": float(np.mean(rep.abs_err_net <= 1e-3)),
                                   "median": float(np.median(rep.abs_err_net)),
                                   "max": float(rep.abs_err_net.max())},
        "seconds": {"train": rep.train_s, "infer": rep.infer_s, "fallback": rep.fallback_s,
                    "surrogate_query": rep.surrogate_s, "plain_query": rep.plain_s},
        "plain_vs_shipped_grid_max_abs_diff": rep.grid_max_diff,
        "train_loss": rep.train_loss,
        "std_bound_tradeoff": rep.tradeoff,
    }


def _term_fingerprint(x, z, c) -> dict:
    """Order-independent fingerprint of a term set: count and two wrap-around
    u64 sums of a 64-bit mix of each key (alone, and with its coefficient bits)."""
    def mix(v):
        v = (v ^ (v >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        v = (v ^ (v >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return v ^ (v >> np.uint64(31))
    with np.errstate(over="ignore"):
        h = np.full(len(c), np.uint64(0x9E3779B97F4A7C15))
        for col in list(np.asarray(x, dtype=np.uint64).T) + list(np.asarray(z, dtype=np.uint64).T):
            h = mix(h ^ col)
        hc = mix(h ^ np.ascontiguousarray(c, dtype=np.float64).view(np.uint64))
        return {"terms": int(len(c)), "key_sum": int(h.sum(dtype=np.uint64)),
                "key_coef_sum": int(hc.sum(dtype=np.uint64))}


def run_parity(args) -> dict:
    """Full-size parity of the headline workload: the whole 20-step 10x10
    lattice propagation (52M live terms) on the GPU against the oracle's
    sharded multi-thread variant on the host cores (same term set and
    coefficient bits; the sharded variant stores terms in another order, so
    the comparison is by order-independent fingerprints), plus the full-size
    CPU rate."""
    from oracle.pyoracle import OracleSum
    from ppvm_amd import workloads as WL
    w = WL.lattice_tfim()
    t0 = time.perf_counter()
    g = _make(w)
    g.apply(w.program)
    g.synchronize()
    gpu_s = time.perf_counter() - t0
    gx, gz, gc = g._export()
    g_ovl = g.overlap_with_zero()
    g_u = g._h.counters()["term_updates"]
    del g
    fg = _term_fingerprint(gx, gz, gc)
    del gx, gz, gc
    _log(f"gpu: {gpu_s:.2f}s, {fg['terms']} terms")
    threads = max(1, min(16, os.cpu_count() or 1))
    o = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
    o.add_strings([s for s, _ in w.observable], [c for _, c in w.observable])
    t0 = time.perf_counter()
    prev = 0
    for li, m in enumerate(w.layer_marks):  # one call per layer, so progress shows
        o.apply_sharded(w.program[prev:m + 1], threads)
        prev = m + 1
        if li % 3 == 2:
            _log(f"cpu sharded: step {li // 3 + 1} done, {len(o)} terms, {time.perf_counter() - t0:.1f}s")
    cpu_s = time.perf_counter() - t0
    ox, oz, oc = o.export()
    fo = _term_fingerprint(ox, oz, oc)
    o_ovl = float(np.sum(oc[~np.any(ox != 0, axis=1)]))
    _log(f"cpu sharded x{threads}: {cpu_s:.1f}s, {fo['terms']} terms")
    return {
        "metric": "full-size parity: 10x10 lattice TFIM, GPU vs sharded CPU oracle",
        "value": 1.0 if fg == fo else 0.0,
        "unit": "identical term set and coefficient bits (1 = yes)",
        "n_gpus": 1,
        "config": {"workload": "lattice_tfim: 10x10 open TFIM, J=1, h=3.04, dt=0.05, 20 Trotter steps, "
                               "Z(5,5), threshold 1e-5, max weight 10, truncate after every gate"},
        "gpu": {"seconds": gpu_s, "fingerprint": fg, "overlap_with_zero": g_ovl, "term_updates": g_u},
        "cpu_sharded": {"seconds": cpu_s, "threads": threads, "fingerprint": fo,
                        "overlap_with_zero": o_ovl, "term_updates": o.term_updates(),
                        "rate": o.term_updates() / cpu_s, "cpu_model": cpu_model()},
        "overlap_abs_diff": abs(g_ovl - o_ovl),
    }


def run_lossy(args) -> dict:
    """Lossy sums: the 10x10 lattice TFIM on lossy words (loss channel on every
    qubit after each Trotter step, reset-loss at readout, max loss weight 2) on
    the GPU lossy engine, checked bit for bit against the oracle in the same run;
    term-updates counted by the oracle (identical term sets), CPU rate from the
    oracle's single-thread run."""
    from oracle.pyoracle import LossyOracleSum
    from ppvm_amd import workloads as WL
    from ppvm_amd._native import LossyHandle
    from ppvm_amd.ops import pack_lossy_strings
    w = WL.lossy_lattice_tfim()
    mlw = w.params["max_loss_weight"]
    obs_x, obs_z, obs_l = pack_lossy_strings([s for s, _ in w.observable], w.n_qubits)
    obs_c = np.array([c for _, c in w.observable])

    def fresh():
        h = LossyHandle(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight, mlw, 1 << 22)
        h.add_words(obs_x, obs_z, obs_l, obs_c)
        return h

    for i in range(args.warmup):
        g = fresh()
        g.apply(w.program)
        _log(f"lossy warmup {i}: {len(g)} terms")
    times = []
    for i in range(args.steps):
        g = fresh()
        t0 = time.perf_counter()
        g.apply(w.program)
        times.append(time.perf_counter() - t0)
        _log(f"lossy run {i}: {times[-1]:.3f}s, {len(g)} terms")
    gpu_s = min(times)
    gx, gz, gl, gc = g.export()
    g_ovl = g.overlap_with_zero()
    res = {
        "metric": "term-updates/s, lossy 10x10 lattice TFIM (lossy engine)",
        "unit": "term-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": gpu_s * 1e3, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "lossy_lattice_tfim: 10x10 open TFIM, J=1, h=3.04, dt=0.05, 4 Trotter steps, "
                               "loss p=0.01 on every qubit per step, reset-loss at readout, Z(5,5), "
                               "threshold 1e-5, max weight 10, max loss weight 2, truncate after every op"},
        "gpu": {"seconds": gpu_s, "seconds_all": times, "terms": int(len(gc)), "overlap_with_zero": g_ovl},
    }
    if args.no_cpu:
        res["value"] = None
        return res
    o = LossyOracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight, mlw)
    o.add_lossy_words(obs_x, obs_z, obs_l, obs_c)
    t0 = time.perf_counter()
    prev = 0
    for li, m in enumerate(w.layer_marks):
        o.apply(w.program[prev:m + 1])
        prev = m + 1
        _log(f"lossy oracle: layer {li + 1}/{len(w.layer_marks)}, {len(o)} terms, {time.perf_counter() - t0:.1f}s")
    o.apply(w.program[prev:])
    cpu_s = time.perf_counter() - t0
    ox, oz, ol, oc = o.export()
    same = (len(oc) == len(gc) and np.array_equal(ox, gx) and np.array_equal(oz, gz) and np.array_equal(ol, gl)
            and np.array_equal(oc.view(np.uint64), gc.view(np.uint64)))
    u = o.term_updates()
    res.update({
        "value": u / gpu_s,
        "term_updates": u,
        "parity": {"identical_terms_order_and_bits": bool(same), "overlap_abs_diff": abs(g_ovl - o.overlap_with_zero())},
        "cpu_baseline": {"value": u / cpu_s, "unit": "term-updates/s", "cores": 1, "kind": "port",
                         "sample": "the whole lossy workload on the oracle", "seconds": cpu_s,
                         "cpu_model": cpu_model()},
        "speedup_vs_cpu": cpu_s / gpu_s,
    })
    return res


def run_configs(args) -> dict:
    """The other single-sum configurations, each a full backward propagation:
    the 20-qubit chain TFIM and the 127-qubit random Clifford + rz circuit.
    GPU time (best of --steps runs on a handle created before the clock), term
    updates, the single-thread oracle's time on this host, and bit-identity of
    the final terms."""
    from ppvm_amd import workloads as WL
    rows = []
    for w in (WL.chain_tfim(), WL.random_circuit()):
        times = []
        for i in range(args.warmup + args.steps):
            g = _make(w, 1 << 18)
            t0 = time.perf_counter()
            g.apply(w.program)
            g.synchronize()
            if i >= args.warmup:
                times.append(time.perf_counter() - t0)
        gx, gz, gc = g._export()
        u = g._h.counters()["term_updates"]
        gpu_s = min(times)
        row = {"workload": w.name, "n_qubits": w.n_qubits, "gates": w.n_gates, "gpu_ms": gpu_s * 1e3,
               "term_updates": u, "value": u / gpu_s, "peak_live": g.stats()["peak_live"],
               "final_terms": int(len(gc)), "overlap_with_zero": g.overlap_with_zero()}
        if not args.no_cpu:
            from oracle.pyoracle import OracleSum
            o = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
            o.add_strings([t for t, _ in w.observable], [c for _, c in w.observable])
            t0 = time.perf_counter()
            o.apply(w.program)
            cpu_s = time.perf_counter() - t0
            ox, oz, oc = o.export()
            row.update({"oracle_ms": cpu_s * 1e3, "cpu_value": o.term_updates() / cpu_s,
                        "identical": bool(np.array_equal(ox, gx) and np.array_equal(oz, gz)
                                          and np.array_equal(oc.view(np.uint64), gc.view(np.uint64)))})
        _log(f"config {w.name}: {row}")
        rows.append(row)
    return {"metric": "term-updates/s per single-sum configuration", "unit": "term-updates/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f64", "data": "synthetic", "rows": rows,
            "config": {"workload": "chain_tfim (n=20, J=1, h=1, dt=0.1, 10 steps, Z_10, threshold 1e-4); "
                                   "random_circuit (127 qubits, 30 layers, threshold 1e-4, max weight 12)"}}


# Wall times recorded in crates/ppvm-pauli-sum/examples/trotter_storage_cliff.rs:16-20
# for this circuit through the reference's Python API (hardware not stated).
REFERENCE_EXAMPLE_MS = {32: 223.0, 40: 544.0, 48: 1044.0, 56: 1842.0, 64: 401.0}


def run_example(args) -> dict:
    """The reference's own timed example (examples/trotter_storage_cliff.rs):
    noisy 1D TFIM, N = 32..64, 20 steps, threshold 1e-6, truncation after every
    call.  GPU time per propagation (best of --steps runs, handle created and
    loaded before the clock), the oracle's single-thread time on this host,
    bit-identity of the final terms, and the reference's recorded times."""
    from oracle.pyoracle import OracleSum
    from ppvm_amd import workloads as WL
    rows = []
    for n in (32, 40, 48, 56, 64):
        w = WL.noisy_chain_tfim(n)
        times = []
        for i in range(args.warmup + args.steps):
            g = _make(w, 1 << 18)
            t0 = time.perf_counter()
            g.apply(w.program)
            g.synchronize()
            if i >= args.warmup:
                times.append(time.perf_counter() - t0)
        gx, gz, gc = g._export()
        u = g._h.counters()["term_updates"]
        row = {"n": n, "gpu_ms": min(times) * 1e3, "terms": int(len(gc)), "term_updates": u,
               "reference_ms_recorded": REFERENCE_EXAMPLE_MS[n]}
        if not args.no_cpu:
            o = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
            o.add_strings([t for t, _ in w.observable], [c for _, c in w.observable])
            t0 = time.perf_counter()
            o.apply(w.program)
            row["oracle_ms"] = (time.perf_counter() - t0) * 1e3
            ox, oz, oc = o.export()
            row["identical"] = bool(np.array_equal(ox, gx) and np.array_equal(oz, gz)
                                    and np.array_equal(oc.view(np.uint64), gc.view(np.uint64)))
        _log(f"example n={n}: {row}")
        rows.append(row)
    return {
        "metric": "ms per propagation, reference example trotter_storage_cliff (noisy 1D TFIM)",
        "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "noisy_chain_tfim: sum_i Z_i, J=1, h=1, dt=0.1, 20 steps, pauli_error 2.5e-5 "
                               "x3 around every gate, threshold 1e-6, truncate after every call"},
        "rows": rows,
        "note": "reference_ms_recorded: wall times written in the reference example's comments "
                "(Python API, hardware not stated)",
    }


# ---- HBM traffic counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) ----------------------
CAL_ELEMENTS = 1 << 28


def run_calibrate(args) -> dict:
    """Calibration workload for the traffic counters: one elementwise kernel of
    known traffic (b = a * 1 over 2^28 float32: 1 GiB read, 1 GiB written),
    three times.  Its FETCH_SIZE / WRITE_SIZE readings per dispatch, divided
    into 1 GiB, are the factors that turn the same counters into bytes."""
    import torch
    a = torch.ones(CAL_ELEMENTS, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    assert float(b[-1]) == 1.0
    return {"metric": "counter calibration", "kernel": "elementwise mul (b = a * 1)", "elements": CAL_ELEMENTS,
            "bytes_read": 4 * CAL_ELEMENTS, "bytes_written": 4 * CAL_ELEMENTS, "dispatches": 3}


def run_gather_ceiling(args) -> dict:
    """What a bare gather of the rotation gather's access pattern achieves on
    this GPU: 52.3M arena entries (32-byte keys, 8-byte coefficients: the
    lattice workload's last arena at W = 2), 8.1% of them selected uniformly at
    random (the anticommuting share of its last gates), their keys and
    coefficients read and written densely (torch index_select, one kernel per
    array).  Reported as algorithmic bytes (selected bytes read + written) per
    second, the same measure the rotation gather's roofline uses."""
    import torch
    n, sel_p, reps = 52_307_346, 0.081, 10
    g = torch.Generator(device="cuda")
    g.manual_seed(12345)
    keys = torch.randint(0, 2**62, (n, 4), device="cuda", dtype=torch.int64, generator=g)
    coef = torch.rand(n, device="cuda", dtype=torch.float64, generator=g)
    idx = torch.nonzero(torch.rand(n, device="cuda", generator=g) < sel_p).squeeze(1)
    m = int(idx.numel())
    for _ in range(2):
        ok, oc = keys.index_select(0, idx), coef.index_select(0, idx)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        ok, oc = keys.index_select(0, idx), coef.index_select(0, idx)
    t1.record()
    torch.cuda.synchronize()
    sec = t0.elapsed_time(t1) / 1e3 / reps
    assert torch.equal(ok[-1], keys[idx[-1]]) and float(oc[0]) == float(coef[idx[0]])
    alg = m * (32 + 8) * 2 + m * 8  # selected keys and coefficients read and written, indices read
    return {"metric": "bare gather of the rotation gather's access pattern", "entries": n, "selected": m,
            "us_per_gather": sec * 1e6, "alg_bytes": alg, "alg_GBs": alg / sec / 1e9,
            "frac_of_8TBs": alg / sec / 8e12}


def _short_kernel(name: str) -> str:
    import re
    m = re.search(r"\b(k_[A-Za-z0-9_]+)", name)
    return m.group(1) if m else name.split("(")[0][-80:]


def _last_gates(rows: list, last_gates: int) -> list:
    """The dispatches (sorted rows whose second field is the short kernel name)
    of the last `last_gates` rotation gates: from the k_rot_count before the
    last_gates-th last k_rot_gather on."""
    gi = [i for i, r in enumerate(rows) if r[1] == "k_rot_gather"]
    start = gi[-last_gates] if len(gi) >= last_gates else 0
    while start > 0 and rows[start][1] != "k_rot_count":
        start -= 1
    return rows[start:]


def counter_rows(root: str, counter: str, last_gates: int = 0) -> dict:
    """kernel -> (dispatches, mean counter value) from the counter_collection
    CSVs rocprofv3 wrote under root; with last_gates > 0 only the dispatches of
    the last `last_gates` rotation gates (from the k_rot_count before the
    last_gates-th last k_rot_gather on)."""
    import csv
    import glob
    from collections import defaultdict
    rows = []
    for f in sorted(glob.glob(os.path.join(root, "**", "*counter_collection.csv"), recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((int(r["Dispatch_Id"]), _short_kernel(r["Kernel_Name"]), float(r["Counter_Value"])))
    rows.sort()
    if last_gates:
        rows = _last_gates(rows, last_gates)
    acc = defaultdict(lambda: [0, 0.0])
    for _, k, v in rows:
        acc[k][0] += 1
        acc[k][1] += v
    return {k: (n, tot / n) for k, (n, tot) in acc.items()}


def run_traffic(args) -> dict:
    """HBM bytes per dispatch from four rocprofv3 --pmc runs (calibration and
    workload, FETCH_SIZE and WRITE_SIZE): each counter scaled by 1 GiB over
    what it read for the calibration kernel."""
    cf = counter_rows(args.cal_fetch, "FETCH_SIZE")
    cw = counter_rows(args.cal_write, "WRITE_SIZE")
    cal = next(k for k in cf if "MulFunctor" in k or "mul" in k.lower())
    kf = (4 * CAL_ELEMENTS / 1024.0) / cf[cal][1]  # KiB per counter unit
    kw = (4 * CAL_ELEMENTS / 1024.0) / cw[cal][1]
    f = counter_rows(args.fetch, "FETCH_SIZE", args.last_gates)
    w = counter_rows(args.write, "WRITE_SIZE", args.last_gates)
    out = {}
    for k in sorted(f, key=lambda k: -f[k][1] * f[k][0]):
        if not k.startswith("k_") or k not in w:
            continue
        out[k] = {"dispatches": f[k][0], "fetch_KiB": f[k][1], "write_KiB": w[k][1],
                  "hbm_bytes": (f[k][1] * kf + w[k][1] * kw) * 1024}
    return {"metric": "HBM bytes per dispatch (rocprofv3 FETCH_SIZE, WRITE_SIZE, calibrated)",
            "fetch_factor": kf, "write_factor": kw, "last_gates": args.last_gates, "kernels": out}


def run_kstats(args) -> dict:
    """Per-kernel statistics from a rocprofv3 --kernel-trace --stats run
    (the *kernel_stats.csv under --dir), durations in microseconds; with
    --last-gates N, from the kernel trace of the last N rotation gates only."""
    import csv
    import glob
    if args.last_gates:
        tr = []
        for f in sorted(glob.glob(os.path.join(args.dir, "**", "*kernel_trace.csv"), recursive=True)):
            with open(f, newline="") as fh:
                for r in csv.DictReader(fh):
                    tr.append((int(r["Start_Timestamp"]), _short_kernel(r["Kernel_Name"]),
                               int(r["End_Timestamp"])))
        tr.sort()
        tr = _last_gates(tr, args.last_gates)
        acc = {}
        for t0, k, t1 in tr:
            d = (t1 - t0) / 1e3
            a = acc.setdefault(k, [0, 0.0, float("inf"), 0.0])
            a[0] += 1
            a[1] += d
            a[2] = min(a[2], d)
            a[3] = max(a[3], d)
        busy = sum(a[1] for a in acc.values())
        rows = [{"kernel": k, "calls": a[0], "total_ms": a[1] / 1e3, "avg_us": a[1] / a[0],
                 "pct": 100.0 * a[1] / busy, "min_us": a[2], "max_us": a[3]} for k, a in acc.items()]
        rows.sort(key=lambda r: -r["total_ms"])
        return {"metric": "rocprofv3 kernel statistics", "kernels": rows, "last_gates": args.last_gates,
                "span_us": (tr[-1][2] - tr[0][0]) / 1e3 if tr else 0.0, "busy_us": busy}
    rows = []
    for f in sorted(glob.glob(os.path.join(args.dir, "**", "*kernel_stats.csv"), recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                name = r["Name"].replace("(anonymous namespace)", "")
                k = _short_kernel(name)
                if "<" in name.split("(")[0] and k.startswith("k_"):
                    k += "<" + name.split("(")[0].split("<", 1)[1]
                rows.append({"kernel": k, "calls": int(r["Calls"]), "total_ms": float(r["TotalDurationNs"]) / 1e6,
                             "avg_us": float(r["AverageNs"]) / 1e3, "pct": float(r["Percentage"]),
                             "min_us": float(r["MinNs"]) / 1e3, "max_us": float(r["MaxNs"]) / 1e3})
    rows.sort(key=lambda r: -r["total_ms"])
    return {"metric": "rocprofv3 kernel statistics", "kernels": rows}


def print_table(res: dict) -> None:
    """Text table of a kstats or traffic result (what profiles/ holds)."""
    if res["metric"].startswith("rocprofv3 kernel"):
        if res.get("last_gates"):
            print(f"# last {res['last_gates']} rotation gates: span {res['span_us']:.0f} us, "
                  f"kernels busy {res['busy_us']:.0f} us")
        print(f"{'kernel':<26} {'calls':>7} {'total_ms':>10} {'avg_us':>9} {'pct':>6} {'min_us':>8} {'max_us':>9}")
        for r in res["kernels"]:
            print(f"{r['kernel']:<26} {r['calls']:>7d} {r['total_ms']:>10.2f} {r['avg_us']:>9.2f} {r['pct']:>6.2f} "
                  f"{r['min_us']:>8.2f} {r['max_us']:>9.2f}")
        return
    print(f"# HBM traffic per dispatch: FETCH_SIZE x {res['fetch_factor']:.4f} + WRITE_SIZE x "
          f"{res['write_factor']:.4f} (KiB -> bytes x 1024)" +
          (f"; last {res['last_gates']} rotation gates only" if res["last_gates"] else ""))
    print(f"{'kernel':<24} {'dispatches':>10} {'fetch_KiB':>12} {'write_KiB':>12} {'hbm_bytes':>14}")
    for k, v in res["kernels"].items():
        print(f"{k:<24} {v['dispatches']:>10d} {v['fetch_KiB']:>12.1f} {v['write_KiB']:>12.1f} {v['hbm_bytes']:>14.0f}")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--mode", choices=["lattice", "batched", "surrogate", "parity", "lossy", "example", "configs",
                                       "calibrate", "traffic", "kstats", "gather-ceiling"], default="lattice")
    ap.add_argument("--dir", help="kstats: rocprofv3 output dir of a --kernel-trace --stats run")
    ap.add_argument("--table", action="store_true", help="traffic / kstats: print a text table, not JSON")
    ap.add_argument("--cal-fetch", help="traffic: rocprofv3 output dir of the calibration FETCH_SIZE run")
    ap.add_argument("--cal-write", help="traffic: ... calibration WRITE_SIZE run")
    ap.add_argument("--fetch", help="traffic: ... workload FETCH_SIZE run")
    ap.add_argument("--write", help="traffic: ... workload WRITE_SIZE run")
    ap.add_argument("--last-gates", type=int, default=0,
                    help="traffic, kstats: only the last N rotation gates (kstats reads the kernel trace)")
    ap.add_argument("--verify", type=int, default=16, help="batched: instances re-run on the engine")
    ap.add_argument("--cpu-instances", type=int, default=256, help="batched: at most this many instances on the CPU")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="batched: CPU baseline time budget (s)")
    ap.add_argument("--surrogate-bound", type=float, default=None,
                    help="surrogate: fixed bound on the ensemble std (default: calibrated on held-out points)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-prefix-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding a sum across GPUs is not "
                                   "implemented (--gpus must be 1)", "n_gpus": args.gpus}))
        return 2
    res = {"lattice": run, "batched": run_batched, "surrogate": run_surrogate,
           "parity": run_parity, "lossy": run_lossy, "example": run_example,
           "configs": run_configs, "calibrate": run_calibrate, "traffic": run_traffic,
           "kstats": run_kstats, "gather-ceiling": run_gather_ceiling}[args.mode](args)
    if args.table and args.mode in ("traffic", "kstats"):
        print_table(res)
        return 0
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
