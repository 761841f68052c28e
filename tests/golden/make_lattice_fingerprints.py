"""Writes tests/golden/lattice_full_fingerprints.json: the oracle's state after
every Trotter step of the full 10x10 lattice TFIM (20 steps, 52.3M live terms
at the end) as order-independent digests: live term count, the sum of the key
digests, the sum of the key-and-coefficient-bit digests (the definition of
ppvm_fingerprint in include/ppvm_hip.h) and overlap_with_zero, plus the term
updates so far.

The single-thread oracle would take hours at this size, so the oracle's
sharded variant runs it (identical term sets and coefficient bits to the
single-thread path, tests/test_oracle.py checks that; only the entry order
differs, which the digests ignore).  No reference fixture covers this circuit,
so these fingerprints are the oracle's own; the oracle itself is pinned by
tests/test_oracle.py.

    python tests/golden/make_lattice_fingerprints.py [--threads 8] [--steps 20] [--state DIR]

With --state DIR the run is resumable: after every step the rows so far and
the oracle's terms are saved there, and a new run continues from the last
saved step (the term set and coefficient bits do not depend on the order in
which the terms are re-added, since no key receives more than one
contribution besides its own per gate).
"""

import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.pyoracle import OracleSum  # noqa: E402
from ppvm_amd import workloads as WL  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "lattice_full_fingerprints.json"))
    ap.add_argument("--state", default=None, help="directory for resumable progress")
    args = ap.parse_args()
    w = WL.lattice_tfim()
    o = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
    rows = []
    prev = 0
    tu0 = 0
    rows_path = os.path.join(args.state, "rows.json") if args.state else None
    if rows_path and os.path.exists(rows_path):
        with open(rows_path) as f:
            rows = json.load(f)
        st = np.load(os.path.join(args.state, "terms.npz"))
        o.add_words(st["x"], st["z"], st["c"])
        prev = rows[-1]["ops"]
        tu0 = rows[-1]["term_updates"]
        assert o.fingerprint() == (rows[-1]["terms"], int(rows[-1]["key_digest"]), int(rows[-1]["key_coef_digest"]))
        print(f"resumed after step {len(rows)}", flush=True)
    else:
        o.add_strings([s for s, _ in w.observable], [c for _, c in w.observable])
    t0 = time.time()
    marks = w.layer_marks
    for step in range(len(rows), args.steps):
        end = marks[3 * step + 2] + 1  # rx layer, then the two rzz layers, per backward step
        o.apply_sharded(w.program[prev:end], args.threads)
        prev = end
        n, ks, kcs = o.fingerprint()
        rows.append({"step": step + 1, "ops": end, "terms": n, "key_digest": str(ks), "key_coef_digest": str(kcs),
                     "overlap": o.overlap_with_zero(), "term_updates": tu0 + o.term_updates()})
        print(f"step {step + 1}: {n} terms, {time.time() - t0:.0f}s", flush=True)
        if args.state and step + 1 < args.steps:
            os.makedirs(args.state, exist_ok=True)
            x, z, c = o.export()
            np.savez(os.path.join(args.state, "terms.tmp.npz"), x=x, z=z, c=c)
            os.replace(os.path.join(args.state, "terms.tmp.npz"), os.path.join(args.state, "terms.npz"))
            with open(rows_path + ".tmp", "w") as f:
                json.dump(rows, f)
            os.replace(rows_path + ".tmp", rows_path)
    out = {"workload": "lattice_tfim: 10x10 open TFIM, J=1, h=3.04, dt=0.05, 20 Trotter steps, Z(5,5), "
                       "threshold 1e-5, max weight 10, truncate after every gate",
           "digest": "ppvm_fingerprint (include/ppvm_hip.h); digests as decimal strings", "steps": rows}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
