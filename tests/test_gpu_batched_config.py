"""The batched workload as specified: all 256 seed-21 TFIM chains
(L~U{32..128}, h~U[0.5,1.5], steps~U{5..30}, threshold 1e-5) through the
batched mode (persistent kernel, with instances that outgrow the arena cap on
the full-GPU engine), against the oracle's per-layer term counts and
overlaps, term updates and final term-set digests in
tests/golden/batched_chains_marks.json (tests/golden/make_batched_marks.py).
"""

import json
import os

import numpy as np
import pytest

from ppvm_amd import workloads as WL
from tests.helpers import host_fingerprint

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batched_chains_marks.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)["instances"]


def test_fixture_covers_the_workload():
    ws = WL.batched_chains()
    g = _golden()
    assert len(g) == 256 == len(ws)
    for i, (w, r) in enumerate(zip(ws, g)):
        assert r["index"] == i and r["n_qubits"] == w.n_qubits and r["layers"] == len(w.layer_marks)


@pytest.mark.gpu
@pytest.mark.parametrize("cap", [1 << 16, 1 << 18])
def test_all_256_instances_match_the_oracle(cap):
    from ppvm_amd.batch import Batch, run_hybrid
    ws = WL.batched_chains()
    g = _golden()
    b = Batch(ws, capacity=cap)
    r = run_hybrid(b, threads=256, use_graph=True, engine_threads=8)
    engine = set(r.on_engine)
    print(f"cap {cap}: {len(ws) - len(engine)} instances in the persistent kernel, {len(engine)} on the engine",
          flush=True)
    assert len(engine) < len(ws)
    for i, w in enumerate(ws):
        want = g[i]
        assert np.array_equal(np.asarray(r.lens[i], dtype=np.uint64), np.asarray(want["lens"], dtype=np.uint64)), \
            f"instance {i}: layer term counts differ"
        assert float(np.max(np.abs(np.asarray(r.overlaps[i]) - np.asarray(want["overlaps"])))) <= 1e-10, \
            f"instance {i}: overlaps differ"
        assert int(r.term_updates[i]) == want["term_updates"], f"instance {i}: term updates"
        fp = r.digests[i] if i in engine else host_fingerprint(*b.export(i))
        assert (fp[0], str(fp[1]), str(fp[2])) == (want["final"][0], want["final"][1], want["final"][2]), \
            f"instance {i}: final terms differ"
