"""Writes tests/golden/batched_chains_marks.json: for each of the 256 batched
TFIM chains of the batched workload (L~U{32..128}, h~U[0.5,1.5],
steps~U{5..30}, seed 21, threshold 1e-5), the oracle's live term count and
overlap_with_zero after every layer, its final fingerprint
(ppvm_fingerprint's digest: count, key digest sum, key+coefficient digest
sum) and its term updates.  No reference fixture covers these circuits, so
the values are the oracle's own (the oracle is pinned by tests/test_oracle.py).

    python tests/golden/make_batched_marks.py [--workers 8]
"""

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def one(i):
    from oracle.pyoracle import OracleSum
    from ppvm_amd import workloads as WL
    w = WL.batched_chains()[i]
    o = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
    o.add_strings([s for s, _ in w.observable], [c for _, c in w.observable])
    t0 = time.time()
    lens, ovl = o.run_marked(w.program, w.layer_marks)
    n, ks, kcs = o.fingerprint()
    return {"index": i, "n_qubits": w.n_qubits, "layers": len(lens), "lens": [int(v) for v in lens],
            "overlaps": [float(v) for v in ovl], "final": [n, str(ks), str(kcs)],
            "term_updates": o.term_updates(), "seconds": time.time() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    from ppvm_amd import workloads as WL
    ws = WL.batched_chains()
    # biggest instances first (most gates) so the pool stays busy
    order = sorted(range(len(ws)), key=lambda i: -len(ws[i].program) * ws[i].n_qubits)
    rows = {}
    with ProcessPoolExecutor(args.workers) as pool:
        for r in pool.map(one, order):
            rows[r["index"]] = r
            print(f"instance {r['index']}: {r['lens'][-1] if r['lens'] else 0} terms, "
                  f"peak {max(r['lens']) if r['lens'] else 0}, {r['seconds']:.1f}s", flush=True)
    out = {"workload": "256 1D TFIM chains, L~U{32..128}, h~U[0.5,1.5], J=1, dt=0.1, steps~U{5..30} (seed 21), "
                       "Z at L//2, threshold 1e-5, marks after every layer",
           "instances": [rows[i] for i in range(len(ws))]}
    for r in out["instances"]:
        del r["seconds"]
    with open(os.path.join(ROOT, "tests", "golden", "batched_chains_marks.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
