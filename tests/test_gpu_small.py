"""Small sums held in LDS (k_small) against the oracle.

A handle whose synchronised arena is small collects the planner's steps and
runs them in one workgroup with the sum in LDS, handing over to the arena
pipeline when the sum outgrows LDS.  Bar: the oracle's terms, order and
coefficient bits, and its term-update count, on the chain TFIM and the
127-qubit random circuit (both stay in LDS), on a lattice prefix that
outgrows LDS, with amplitude damping and noise, at every workgroup size.
"""

import numpy as np
import pytest

from ppvm_amd import ops as O
from ppvm_amd import workloads as WL
from tests.helpers import assert_same_terms, gpu_for, oracle_for, random_program, random_terms, run_segments

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _defaults(monkeypatch):
    for k in ("PPVM_SMALL_MAX_ENTRIES", "PPVM_SMALL_THREADS", "PPVM_PARTITION_MIN_ENTRIES"):
        monkeypatch.delenv(k, raising=False)


@pytest.mark.parametrize("wname", ["chain_tfim", "random_circuit"])
@pytest.mark.parametrize("threads", ["64", "256", "1024"])
def test_small_configs_run_in_lds(wname, threads, monkeypatch):
    monkeypatch.setenv("PPVM_SMALL_THREADS", threads)
    w = getattr(WL, wname)()
    g, r = gpu_for(w), oracle_for(w)
    run_segments(g, r, w.program, w.layer_marks, ctx=f"{wname} threads {threads}")
    c = g._h.counters()
    assert c["term_updates"] == r.term_updates()
    assert c["small_runs"] >= 1 and c["small_steps"] > 0
    assert c["map_launches"] == 0 and c["transposed"] == 0  # no arena-pipeline kernel ran


@pytest.mark.parametrize("whole", [False, True])
def test_small_hands_over_when_the_sum_outgrows_lds(whole):
    """Four lattice steps (3,588 terms), applied layer by layer or as one
    program (one run from a single term, handing over mid-program)."""
    w = WL.truncate_steps(WL.lattice_tfim(), 4)
    g, r = gpu_for(w), oracle_for(w)
    marks = [len(w.program) - 1] if whole else w.layer_marks
    run_segments(g, r, w.program, marks, ctx=f"lattice 4 steps whole {whole}")
    c = g._h.counters()
    assert c["term_updates"] == r.term_updates()
    assert c["small_runs"] >= 1 and c["transposed"] > 0  # started in LDS, finished in the arena


def test_small_with_damping_and_noise():
    rng = np.random.default_rng(77)
    n = 9
    strings = random_terms(rng, n, 30)
    coeffs = rng.normal(size=30)
    from oracle.pyoracle import OracleSum
    from ppvm_amd import PauliSum
    g = PauliSum(strings, n, coefficients=coeffs, min_abs_coeff=1e-4, max_pauli_weight=6)
    r = OracleSum(n, 1e-4, 6)
    r.add_strings(strings, coeffs)
    p = O.Program()
    for i in range(60):
        a, b = (int(v) for v in rng.choice(n, size=2, replace=False))
        p.add(O.RX, a, theta=float(rng.uniform(-1, 1))).truncate()
        p.add(O.CNOT, a, b)
        if i % 3 == 0:
            p.add(O.AMPLITUDE_DAMPING, b, p=[0.1]).truncate()
        if i % 4 == 1:
            p.add(O.DEPOLARIZE1, a, p=[0.02]).truncate()
        if i % 5 == 2:
            p.add(O.RZZ, a, b, theta=0.3)
    prog = p.array()
    run_segments(g, r, prog, list(range(10, len(prog), 25)), ctx="damping")
    assert g._h.counters()["small_runs"] >= 1


@pytest.mark.parametrize("n", [40, 100, 200, 400])
def test_small_every_word_count(n):
    """W = 1, 2, 4, 8 with truncation after every gate and a weight cap."""
    rng = np.random.default_rng(300 + n)
    strings = random_terms(rng, n, 8)
    coeffs = rng.normal(size=8)
    from oracle.pyoracle import OracleSum
    from ppvm_amd import PauliSum
    g = PauliSum(strings, n, coefficients=coeffs, min_abs_coeff=1e-3, max_pauli_weight=n // 2 + 5)
    r = OracleSum(n, 1e-3, n // 2 + 5)
    r.add_strings(strings, coeffs)
    prog = random_program(rng, n, 200, trunc_every=True)
    run_segments(g, r, prog, list(range(7, len(prog), 40)), ctx=f"n={n}")
    assert_same_terms(g._export(), r.export(), True, f"n={n} end")
    assert g._h.counters()["small_runs"] >= 1


def test_hand_over_launches_no_rotation_that_acts_on_no_term(monkeypatch):
    """The steps an LDS run leaves when it outgrows LDS were planned with the
    support widened over the whole collection; the hand-over re-decides them
    from the exact support of the state the run left.  A lattice prefix that
    starts in LDS therefore launches no more rotation pipelines than the same
    run on the arena pipeline alone (where the support is exact at every
    synchronisation), with the same results."""
    w = WL.truncate_steps(WL.lattice_tfim(), 8)

    def launched(small):
        if small:
            monkeypatch.delenv("PPVM_SMALL_MAX_ENTRIES", raising=False)
        else:
            monkeypatch.setenv("PPVM_SMALL_MAX_ENTRIES", "0")
        g = gpu_for(w)
        g._h.profile(True)
        g.apply(w.program)
        g.synchronize()
        _, n = g._h.profile_read()
        return int(n[0]), g._h.counters(), g._h.fingerprint()

    a, ca, fa = launched(True)
    b, cb, fb = launched(False)
    print(f"rotation pipelines launched: {a} with the LDS start, {b} without", flush=True)
    assert ca["small_runs"] >= 1 and cb["small_runs"] == 0
    assert fa == fb and ca["term_updates"] == cb["term_updates"]
    assert 0 < a <= b
