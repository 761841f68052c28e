"""Batched persistent-kernel mode vs the CPU oracle and the single-instance engine.

Bar: every instance's final terms bit-identical (keys, coefficient bits and
insertion order) to the oracle's; recorded (len, overlap) at every layer equal
to the oracle's (overlap within 1e-10)."""

import numpy as np
import pytest

from ppvm_amd import ops as O
from ppvm_amd import workloads as WL
from tests.helpers import assert_same_terms, oracle_for, random_program, random_terms

pytestmark = pytest.mark.gpu


def _small_chains(count=12, seed=5):
    rng = WL.SplitMix64(seed)
    out = []
    for i in range(count):
        L = rng.randint(8, 70)
        steps = rng.randint(2, 5)
        h = 0.5 + rng.uniform()
        out.append(WL.chain_tfim(n=L, h=h, steps=steps, obs=L // 2, min_abs_coeff=1e-4))
    return out


def _check_batch(b, ws):
    res = b.results()
    for i, w in enumerate(ws):
        assert res["status"][i] == 0, f"instance {i} status {res['status'][i]}"
        r = oracle_for(w)
        lens, ovl = r.run_marked(w.program, w.layer_marks)
        bl, bo = b.marks(i)
        assert np.array_equal(bl, lens), f"instance {i}: layer term counts differ"
        assert np.max(np.abs(bo - ovl)) <= 1e-10, f"instance {i}: overlaps differ"
        assert_same_terms(b.export(i), r.export(), True, f"instance {i}")
        assert int(res["term_updates"][i]) == r.term_updates()


@pytest.mark.parametrize("threads", [64, 256, 1024])
def test_batch_chains_match_oracle(threads):
    from ppvm_amd.batch import Batch
    ws = _small_chains()
    b = Batch(ws)
    b.run(threads=threads)
    _check_batch(b, ws)


def test_batch_graph_and_segments_match():
    from ppvm_amd.batch import Batch
    ws = _small_chains(8, seed=9)
    b = Batch(ws)
    b.run(threads=256, segment_ops=37, use_graph=True)
    _check_batch(b, ws)
    b.run(threads=256, segment_ops=37, use_graph=True)  # graph replay
    _check_batch(b, ws)


def test_batch_mixed_programs_and_widths():
    """Cliffords, every rotation, noise and weight caps; n from 6 to 130 (W up to 4)."""
    from ppvm_amd.batch import Batch
    rng = np.random.default_rng(17)
    ws = []
    for i, n in enumerate([6, 20, 63, 64, 65, 100, 130]):
        strings = random_terms(rng, n, 5)
        prog = random_program(rng, n, 60, trunc_every=True)
        p = O.Program()
        for op in prog:
            p.add(int(op["code"]), int(op["q0"]), int(op["q1"]), float(op["theta"]))
            if rng.uniform() < 0.1:
                p.add(O.DEPOLARIZE1, int(op["q0"]), p=[0.01]).truncate()
        arr = p.array()
        marks = [j for j in range(len(arr)) if arr[j]["code"] == O.TRUNCATE][::7]
        ws.append(WL.Workload(f"mixed{i}", n, 1e-3, 8 if i % 2 else None,
                              [(s, float(c)) for s, c in zip(strings, rng.normal(size=5))], arr, marks))
    b = Batch(ws, capacity=1 << 14)
    b.run(threads=256)
    _check_batch(b, ws)


def test_batch_overflow_then_regrow():
    from ppvm_amd.batch import Batch
    ws = _small_chains(4, seed=3)
    ws.append(WL.truncate_steps(WL.lattice_tfim(), 4))
    b = Batch(ws, capacity=1024)
    b.run(threads=256)
    res = b.results()
    assert res["status"][-1] == 1  # the lattice instance outgrows 1024 terms
    b.set_capacity(len(ws) - 1, 1 << 16)
    b.run(threads=256)
    _check_batch(b, ws)


def test_hybrid_routes_large_instances_to_engine():
    from ppvm_amd.batch import Batch, run_hybrid
    ws = _small_chains(6, seed=11) + [WL.truncate_steps(WL.lattice_tfim(), 5)]
    b = Batch(ws, capacity=2048)
    r = run_hybrid(b, threads=256, use_graph=False, engine_threads=2)
    assert len(ws) - 1 in r.on_engine
    for i, w in enumerate(ws):
        o = oracle_for(w)
        lens, ovl = o.run_marked(w.program, w.layer_marks)
        assert np.array_equal(np.asarray(r.lens[i], dtype=np.uint64), lens)
        assert np.max(np.abs(np.asarray(r.overlaps[i]) - ovl)) <= 1e-10


@pytest.mark.parametrize("limit", ["1", "3"])
def test_batch_table_generation_wrap(monkeypatch, limit):
    """With the generation limit forced down (PPVM_BATCH_GEN_LIMIT), every
    instance's hash table is cleared and its generation restarted after every
    `limit` rotations; results stay bit-identical to the oracle's, across a
    second run that continues from the generations the first left behind."""
    from ppvm_amd.batch import Batch
    monkeypatch.setenv("PPVM_BATCH_GEN_LIMIT", limit)
    ws = _small_chains(6, seed=13)
    b = Batch(ws)
    b.run(threads=256)
    _check_batch(b, ws)
    b.run(threads=256)
    _check_batch(b, ws)
