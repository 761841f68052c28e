"""Partitioned pairing with oversized buckets, vs the CPU oracle.

With PPVM_PARTITION_MAX_LG=0 the radix partition keeps the whole worklist in
one bucket, which takes the branches that ordinary workloads reach only past
tens of millions of anticommuting terms: pairing in several LDS passes over
sub-bins (bucket > 8192 records), the global table for sub-bins that are still
too large (bucket > ~524k records), and the global table for a whole bucket
whose positions exceed 20 bits (> 1,048,575 records).  Every term anticommutes
with the first gate, Z on qubit 0, and half of them have their partner (Z on
qubit 0 flipped) in the sum.  Bar: identical keys, order and coefficient bits.
"""

import numpy as np
import pytest

from ppvm_amd import ops as O
from tests.helpers import assert_same_terms, run_segments

pytestmark = pytest.mark.gpu

N = 70  # two 64-bit words per plane
PASS_MAX = 8192  # records one LDS pass of the pairing kernel holds (half its table)
SUB_BINS = 64


def _words(rng, m):
    """m distinct keys with X or Y on qubit 0, plus the partners of the first m/2."""
    hi_mask = np.uint64((1 << (N - 64)) - 1)
    x = rng.integers(0, 2**63, size=(m, 2), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=(m, 2), dtype=np.uint64)
    z = rng.integers(0, 2**63, size=(m, 2), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=(m, 2), dtype=np.uint64)
    x[:, 1] &= hi_mask
    z[:, 1] &= hi_mask
    x[:, 0] |= np.uint64(1)  # X or Y on qubit 0
    keys = np.unique(np.concatenate([x, z], axis=1), axis=0)
    rng.shuffle(keys)
    x, z = keys[:, :2].copy(), keys[:, 2:].copy()
    px, pz = x[: len(x) // 2].copy(), z[: len(z) // 2].copy()
    pz[:, 0] ^= np.uint64(1)  # partner under Z_0: X <-> Y on qubit 0
    allk = np.concatenate([np.concatenate([x, z], axis=1), np.concatenate([px, pz], axis=1)])
    _, first = np.unique(allk, axis=0, return_index=True)
    allk = allk[np.sort(first)]
    c = rng.normal(size=len(allk))
    return allk[:, :2].copy(), allk[:, 2:].copy(), c


@pytest.mark.parametrize("method", ["hash", "sort"])
@pytest.mark.parametrize("m", [4_000, 12_000, 200_000, 450_000, 720_000])
def test_single_bucket_branches(m, method, monkeypatch):
    """m = 4,000 fits one pass (6,000 records), which the sort-based pairing
    handles by its counting sort; the larger buckets take the shared
    multi-pass and global-table branches."""
    from oracle.pyoracle import OracleSum
    from ppvm_amd import PauliSum
    monkeypatch.setenv("PPVM_PARTITION_MIN_ENTRIES", "0")
    monkeypatch.setenv("PPVM_SMALL_MAX_ENTRIES", "0")
    monkeypatch.setenv("PPVM_PARTITION_MAX_LG", "0")
    if method == "sort":
        monkeypatch.setenv("PPVM_PAIR_METHOD", "sort")
    else:
        monkeypatch.delenv("PPVM_PAIR_METHOD", raising=False)
    rng = np.random.default_rng(m)
    x, z, c = _words(rng, m)
    total = len(c)
    first = "I" * N
    g = PauliSum([first], N, coefficients=[0.5], min_abs_coeff=0.0)
    r = OracleSum(N, 0.0, None)
    r.add_strings([first], [0.5])
    g._h.add_words(x, z, c)
    r.add_words(x, z, c)
    p = O.Program()
    p.add(O.RZ, 0, 0, 0.3)
    p.add(O.RX, 1, 0, 0.7)
    p.add(O.RZZ, 2, 3, -0.4)
    p.add(O.RZ, 0, 0, 1.1)
    prog = p.array()
    print(f"m={m}: {total} records anticommute with the first gate", flush=True)
    run_segments(g, r, prog, list(range(len(prog))), ctx=f"single bucket m={m}")
    fb = g._h.counters()["pair_fallback"]
    if total <= SUB_BINS * PASS_MAX // 2:
        assert fb == 0, f"{fb} records took the global table"
    if total > (1 << 20):
        assert fb >= total, f"only {fb} of {total} records took the global table"
    elif total > SUB_BINS * PASS_MAX:
        assert fb > 0


@pytest.mark.parametrize("method", ["hash", "sort"])
@pytest.mark.parametrize("target", [16, 64])
def test_many_buckets_per_record_scatter(target, method, monkeypatch):
    """With buckets aimed at 16 or 64 records, 300k records spread over 2^14 or
    2^13 buckets: more than the tiled scatter's 2^12, so the scatter stores
    record by record, and each LDS pass holds a small bucket."""
    from oracle.pyoracle import OracleSum
    from ppvm_amd import PauliSum
    monkeypatch.setenv("PPVM_PARTITION_MIN_ENTRIES", "0")
    monkeypatch.setenv("PPVM_SMALL_MAX_ENTRIES", "0")
    monkeypatch.delenv("PPVM_PARTITION_MAX_LG", raising=False)
    monkeypatch.setenv("PPVM_PARTITION_BUCKET_RECORDS", str(target))
    if method == "sort":
        monkeypatch.setenv("PPVM_PAIR_METHOD", "sort")
    else:
        monkeypatch.delenv("PPVM_PAIR_METHOD", raising=False)
    rng = np.random.default_rng(4242 + target)
    x, z, c = _words(rng, 200_000)
    total = len(c)
    assert total // target > 4096  # more buckets than the tiled scatter takes
    first = "I" * N
    g = PauliSum([first], N, coefficients=[0.5], min_abs_coeff=0.0)
    r = OracleSum(N, 0.0, None)
    r.add_strings([first], [0.5])
    g._h.add_words(x, z, c)
    r.add_words(x, z, c)
    p = O.Program()
    p.add(O.RZ, 0, 0, 0.3)
    p.add(O.RX, 1, 0, 0.7)
    p.add(O.RZZ, 2, 3, -0.4)
    p.add(O.RZ, 0, 0, 1.1)
    prog = p.array()
    run_segments(g, r, prog, list(range(len(prog))), ctx=f"{total} records, {target} per bucket, {method}")
    assert g._h.counters()["pair_fallback"] == 0
