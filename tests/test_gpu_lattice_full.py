"""The headline workload at full size on the HIP engine: every Trotter step of
the 10x10 lattice TFIM (20 steps, 52.3M live terms at the end) against the
oracle's digests in tests/golden/lattice_full_fingerprints.json (written by
tests/golden/make_lattice_fingerprints.py).

The engine runs with its default settings from a one-term handle, so this
covers arena growth with replay, the switch to the partitioned pairing at 4M
entries, the end of support refreshes above 16M entries, dead-entry
compaction and 2^28-entry arenas.  Per step: live term count, the digest of
the key set, the digest of keys with coefficient bits (both exact), term
updates (exact) and overlap_with_zero (1e-10).

The digest itself (ppvm_fingerprint) is checked against a numpy restatement
of its definition and against the oracle's on smaller sums.
"""

import json
import os

import numpy as np
import pytest

from oracle.pyoracle import OracleSum
from ppvm_amd import ops as O
from ppvm_amd import workloads as WL
from tests.helpers import gpu_for, host_fingerprint, oracle_for

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lattice_full_fingerprints.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_fixture_shape():
    g = _golden()
    rows = g["steps"]
    assert [r["step"] for r in rows] == list(range(1, 21))
    assert rows[-1]["terms"] == 52307346
    assert rows[11]["terms"] == 2913437  # the 12-step prefix the bench re-checks


def test_fixture_prefix_matches_single_thread_oracle():
    """The sharded oracle that wrote the fixture agrees with the single-thread
    oracle on the first 8 steps (digests are order-independent)."""
    w = WL.lattice_tfim()
    o = oracle_for(w)
    rows = _golden()["steps"]
    for step in range(8):
        o.apply(w.program[(w.layer_marks[3 * step - 1] + 1 if step else 0):w.layer_marks[3 * step + 2] + 1])
        n, ks, kcs = o.fingerprint()
        r = rows[step]
        assert (n, str(ks), str(kcs)) == (r["terms"], r["key_digest"], r["key_coef_digest"])
        assert o.term_updates() == r["term_updates"]
        assert abs(o.overlap_with_zero() - r["overlap"]) <= 1e-12


def test_host_fingerprint_matches_oracle_digest():
    w = WL.chain_tfim()
    o = oracle_for(w)
    o.apply(w.program)
    assert host_fingerprint(*o.export()) == o.fingerprint()


@pytest.mark.gpu
@pytest.mark.parametrize("wname", ["chain_tfim", "random_circuit"])
def test_device_fingerprint_matches_definition(wname):
    w = getattr(WL, wname)()
    g = gpu_for(w)
    o = oracle_for(w)
    for cut in (len(w.program) // 3, len(w.program)):
        prog = w.program[:cut]
        g2 = gpu_for(w)
        g2.apply(prog)
        o2 = oracle_for(w)
        o2.apply(prog)
        fp = g2._h.fingerprint()
        assert fp == host_fingerprint(*g2._export())
        assert fp == o2.fingerprint()
    del g, o


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["hash", "sort"])
def test_lattice_full_run_every_step(method, monkeypatch):
    """Default settings, with the LDS hash table (the default) or the
    sort-based pairing inside the partition's buckets."""
    for k in ("PPVM_PARTITION_MIN_ENTRIES", "PPVM_SMALL_MAX_ENTRIES", "PPVM_PARTITION_MAX_LG"):
        monkeypatch.delenv(k, raising=False)
    if method == "sort":
        monkeypatch.setenv("PPVM_PAIR_METHOD", "sort")
    else:
        monkeypatch.delenv("PPVM_PAIR_METHOD", raising=False)
    w = WL.lattice_tfim()
    rows = _golden()["steps"]
    g = gpu_for(w)  # default settings: grows from a small arena
    prev = 0
    for step, r in enumerate(rows):
        end = w.layer_marks[3 * step + 2] + 1
        assert end == r["ops"]
        g.apply(w.program[prev:end])
        prev = end
        n, ks, kcs = g._h.fingerprint()
        print(f"step {step + 1}: {n} terms", flush=True)
        assert n == r["terms"], f"step {step + 1}: {n} vs {r['terms']} terms"
        assert str(ks) == r["key_digest"], f"step {step + 1}: key sets differ"
        assert str(kcs) == r["key_coef_digest"], f"step {step + 1}: coefficient bits differ"
        assert g._h.counters()["term_updates"] == r["term_updates"], f"step {step + 1}: term updates"
        assert abs(g.overlap_with_zero() - r["overlap"]) <= 1e-10, f"step {step + 1}: overlap"
    c = g._h.counters()
    # the regimes this run is meant to reach
    assert c["capacity"] >= 1 << 26 and c["compactions"] > 0 and c["regrows"] > 0
