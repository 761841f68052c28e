"""ppvm_batch_run1_noalloc (no device allocation, scratch-backed) against the
allocating batched path (ppvm_batch_create + run + results + export) on the
same instance: identical terms, order and coefficient bits, term updates and
last overlap, on two calls in a row, the second with a smaller input; and the
oracle on the same workload.  Run alone and after the other GPU tests.
"""

import numpy as np
import pytest

from ppvm_amd import workloads as WL
from tests.helpers import assert_same_terms, oracle_for

pytestmark = pytest.mark.gpu

CAP = 1 << 14


def _allocating(w):
    from ppvm_amd.batch import Batch
    b = Batch([w], capacity=CAP)
    b.run(threads=256)
    r = b.results()
    assert int(r["status"][0]) == 0
    x, z, c = b.export(0)
    return x, z, c, float(r["overlap"][0]), int(r["term_updates"][0])


def test_noalloc_matches_allocating_path_twice():
    from ppvm_amd.batch import run1_noalloc
    big = WL.chain_tfim(n=40, steps=8)
    small = WL.chain_tfim(n=20, steps=5)
    for w in (big, small):
        print(f"noalloc: {w.name} n={w.n_qubits} ops={len(w.program)}", flush=True)
        x, z, c, ov, tu = run1_noalloc(w, capacity=CAP)
        ax, az, ac, aov, atu = _allocating(w)
        assert_same_terms((x, z, c), (ax, az, ac), ordered=True, ctx=w.name)
        assert tu == atu
        assert np.float64(ov).view(np.uint64) == np.float64(aov).view(np.uint64)
        o = oracle_for(w)
        o.apply(w.program)
        assert_same_terms((x, z, c), o.export(), ordered=True, ctx=w.name + " vs oracle")


def test_noalloc_rejects_runs_beyond_the_scratch():
    from ppvm_amd.batch import run1_noalloc
    w = WL.chain_tfim(n=20, steps=2)
    with pytest.raises(ValueError, match="scratch"):
        run1_noalloc(w, capacity=1 << 22)
    # a run that still fits works afterwards
    x, z, c, ov, tu = run1_noalloc(w, capacity=1 << 12)
    o = oracle_for(w)
    o.apply(w.program)
    assert_same_terms((x, z, c), o.export(), ordered=True, ctx="after rejection")
