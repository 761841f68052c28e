"""Pins the CPU oracle of lossy sums (PauliSum over {I, X, Y, Z, L} words).

  * the reference's word-level Clifford tables for lossy words
    (crates/ppvm-pauli-word/src/loss/clifford.rs tests, committed as
    tests/golden/lossy_clifford_words.json);
  * the known answers of crates/ppvm-pauli-sum/tests/loss.rs and
    ppvm-python/test/test_loss.py (GHZ with loss, correlated loss, reset-loss
    accumulation, loss truncation, rotations on lost qubits);
  * on words without L the lossy oracle equals the plain oracle bit for bit, and
    gates that only touch kept qubits act on a lossy word as on its x/z planes;
  * loss channels on L-free sums are the closed-form scalings (1 - p) and
    (1 - 2 p1 - p0).
"""

import json
import os

import numpy as np
import pytest

from oracle.pyoracle import LossyOracleSum, OracleSum
from ppvm_amd import ops as O

HERE = os.path.dirname(os.path.abspath(__file__))
NAME_TO_CODE = {**O.CLIFFORD_1Q, **O.CLIFFORD_2Q}


def lossy(n, terms, tau=0.0, maxw=None, maxlw=None):
    s = LossyOracleSum(n, tau, maxw, maxlw)
    if isinstance(terms, dict):
        s.add_strings(list(terms), list(terms.values()))
    else:
        s.add_strings(list(terms))
    return s


def run(s, *ops):
    p = O.Program()
    for op in ops:
        p.add(*op[:4], p=op[4] if len(op) > 4 else None)
    s.apply(p.array())
    return s


def test_clifford_word_tables():
    # loss/clifford.rs: every gate on every one-qubit (or two-qubit) lossy word
    vec = json.load(open(os.path.join(HERE, "golden", "lossy_clifford_words.json")))
    assert set(vec) == set(NAME_TO_CODE)
    for gate, spec in vec.items():
        code = NAME_TO_CODE[gate]
        q = spec["qubits"] + [0]
        for word, target in spec["pairs"]:
            s = lossy(len(word), [word])
            run(s, (code, q[0], q[1] if code in O.TWO_QUBIT else 0))
            (got, c), = s.terms()
            assert got == target, (gate, word)
            if "L" in word:
                assert c == 1.0, (gate, word)  # a lost qubit leaves the sign alone


def test_lossy_sum_and_channels_basic():
    # tests/loss.rs:14-115
    s = lossy(2, {"ZI": 1.0, "IZ": 1.0, "LL": 1.0})
    assert len(s) == 3
    for w in ("X", "Y"):
        assert run(lossy(1, [w]), (O.RESET_LOSS, 0)).terms() == [(w, 1.0)]
    for w in ("I", "Z"):
        assert run(lossy(1, [w]), (O.RESET_LOSS, 0)).terms() == [(w, 1.0), ("L", 1.0)]
    assert run(lossy(1, ["L"]), (O.RESET_LOSS, 0)).terms() == [("L", 0.0)]
    for w in "XYIZ":
        assert run(lossy(1, [w]), (O.LOSS, 0, 0, 0.0, [0.2])).terms() == [(w, 0.8)]
    assert run(lossy(1, ["L"]), (O.LOSS, 0, 0, 0.0, [0.2])).terms() == [("L", 1.0), ("I", 0.2)]


def test_single_qubit_loss_known_answer():
    # tests/loss.rs:117-139
    s = lossy(1, ["Z"])
    run(s, (O.RESET_LOSS, 0), (O.X, 0), (O.X, 0), (O.LOSS, 0, 0, 0.0, [0.1]), (O.X, 0))
    assert abs(s.trace("Z?*") + 0.8) < 1e-10


def test_ghz_with_loss_known_answers():
    p = 0.1
    # tests/loss.rs:141-170
    s = lossy(2, ["ZZ"])
    run(s, (O.RESET_LOSS, 0), (O.RESET_LOSS, 1), (O.X, 0), (O.X, 1), (O.X, 0), (O.X, 1),
        (O.LOSS, 0, 0, 0.0, [p]), (O.LOSS, 1, 0, 0.0, [p]), (O.CNOT, 0, 1), (O.H, 0))
    prob = 0.5 + 0.5 * ((1 - p) * (1 - p) - 2 * p * (1 - p) + p * p)
    assert abs(s.overlap_with_zero() - prob) < 1e-10
    # tests/loss.rs:172-197
    s = lossy(2, ["ZZ"])
    run(s, (O.RESET_LOSS, 0), (O.RESET_LOSS, 1), (O.LOSS, 0, 0, 0.0, [p]), (O.LOSS, 1, 0, 0.0, [p]),
        (O.CNOT, 0, 1), (O.LOSS, 0, 0, 0.0, [2 * p]), (O.H, 0))
    prob = 2 * p + (1 - 2 * p) * (0.5 + 0.5 * ((1 - p) * (1 - p) - 2 * p * (1 - p) + p * p))
    assert abs(s.overlap_with_zero() - prob) < 1e-10
    # ppvm-python/test/test_loss.py::test_ghz (truncating after every call)
    s = lossy(2, ["ZZ"], tau=1e-10)
    run(s, (O.RESET_LOSS, 0), (O.TRUNCATE,), (O.LOSS, 0, 0, 0.0, [p]), (O.TRUNCATE,), (O.CNOT, 0, 1),
        (O.TRUNCATE,), (O.H, 0), (O.TRUNCATE,))
    assert abs(s.overlap_with_zero() - (1 - p)) < 1e-9


def test_correlated_loss_known_answers():
    p = [0.1, 0.05, 0.2]
    # tests/loss.rs:300-352
    s = run(lossy(2, ["ZI"]), (O.CORRELATED_LOSS, 0, 1, 0.0, p))
    assert s.terms() == [("ZI", 1.0 - 2.0 * p[1] - p[0])]
    s = run(lossy(2, ["ZL"]), (O.CORRELATED_LOSS, 0, 1, 0.0, p))
    assert s.terms() == [("ZL", 1.0 - p[2]), ("ZI", p[1])]
    s = run(lossy(2, ["LZ"]), (O.CORRELATED_LOSS, 0, 1, 0.0, p))
    assert s.terms() == [("LZ", 1.0 - p[2]), ("IZ", p[1])]
    # three distinct children against one entry: aux is the larger map, so the
    # children come first and LL (unchanged by the channel) is appended after
    # them (sum/data.rs:181-189)
    s = run(lossy(2, ["LL"]), (O.CORRELATED_LOSS, 0, 1, 0.0, p))
    assert s.terms() == [("IL", p[2]), ("LI", p[2]), ("II", p[0]), ("LL", 1.0)]
    # tests/loss.rs:354-400 and test_loss.py::test_correlated_loss_ghz
    for obs, expect in (("ZZ", 1.0), ("XX", 0.9)):
        s = lossy(2, [obs], tau=1e-10)
        run(s, (O.RESET_LOSS, 0), (O.RESET_LOSS, 1), (O.CORRELATED_LOSS, 0, 1, 0.0, [0.1, 0.0, 0.0]),
            (O.CNOT, 0, 1), (O.H, 0))
        assert abs(s.overlap_with_zero() - expect) < 1e-9


def test_correlated_loss_merges_children_before_the_entry():
    """map_insert_multiple (sum/data.rs:234-247): a word's children are summed
    in parent then slot order in the aux map, and that sum is added to the
    existing coefficient once: p + ((c1 + c2) + c3), not ((p + c1) + c2) + c3.
    Here II receives p1 b (from IL), p1 e (from LI) and p0 d (from LL); the
    values are chosen so that the two associations differ in the last bit.
    Four entries against three distinct children: the entries keep their order."""
    pr = [0.3, 0.1, 0.7]
    a, b, d, e = 1e16, 4.0, 1.3, 4.0
    s = run(lossy(2, {"II": a, "IL": b, "LI": e, "LL": d}), (O.CORRELATED_LOSS, 0, 1, 0.0, pr))
    ap = a * (1.0 - pr[1] * 2.0 - pr[0])  # II itself is in the qubit subspace on both
    c1, c2, c3 = b * pr[1], e * pr[1], d * pr[0]
    want_ii = ap + ((c1 + c2) + c3)
    assert want_ii != ((ap + c1) + c2) + c3  # the case tells the two orders apart
    t = s.terms()
    assert [k for k, _ in t] == ["II", "IL", "LI", "LL"]
    got = dict(t)
    assert got["II"] == want_ii
    assert got["IL"] == b * (1.0 - pr[2]) + d * pr[2]
    assert got["LI"] == e * (1.0 - pr[2]) + d * pr[2]
    assert got["LL"] == d


def test_correlated_loss_larger_aux_puts_children_first():
    # two entries, three distinct children (one of them an existing key):
    # the result is aux's order (IL, LI, II), then the entries not among the
    # children (LL), and IL = children's sum + its own value
    pr = [0.2, 0.1, 0.4]
    s = run(lossy(2, {"LL": 2.0, "IL": 0.5}), (O.CORRELATED_LOSS, 0, 1, 0.0, pr))
    t = s.terms()
    assert [k for k, _ in t] == ["IL", "LI", "II", "LL"]
    got = dict(t)
    assert got["IL"] == 2.0 * pr[2] + 0.5 * (1.0 - pr[2])
    assert got["II"] == 2.0 * pr[0] + 0.5 * pr[1]
    assert got["LL"] == 2.0


def test_reset_loss_accumulates_in_order():
    # tests/loss.rs:228-260: I and Z both land on L
    s = run(lossy(1, {"I": 2.0, "Z": 3.0}), (O.RESET_LOSS, 0))
    assert s.terms() == [("I", 2.0), ("Z", 3.0), ("L", 5.0)]
    # an existing L is zeroed first, then receives both children
    s = run(lossy(1, {"Z": 3.0, "L": 7.0, "I": 2.0}), (O.RESET_LOSS, 0))
    assert s.terms() == [("Z", 3.0), ("L", 5.0), ("I", 2.0)]


def test_loss_weight_truncation():
    # tests/loss.rs:199-226
    s = lossy(3, ["ZZZ"], maxlw=2)
    run(s, (O.RESET_LOSS, 0), (O.RESET_LOSS, 1), (O.RESET_LOSS, 2),
        (O.LOSS, 0, 0, 0.0, [0.1]), (O.LOSS, 1, 0, 0.0, [0.1]), (O.LOSS, 2, 0, 0.0, [0.1]))
    n = len(s)
    run(s, (O.TRUNCATE,))
    assert len(s) == n - 1
    assert all(w.count("L") <= 2 for w, _ in s.terms())
    # the Pauli weight counts L (loss/data.rs:23-31)
    s = lossy(4, {"XLZL": 1.0, "XIZI": 1.0}, maxw=3)
    run(s, (O.TRUNCATE,))
    assert s.terms() == [("XIZI", 1.0)]


def test_rotations_on_lost_qubits():
    # tests/loss.rs:262-298: rx on L is a no-op; rxx with one lost qubit is rx on the other
    assert run(lossy(1, ["L"]), (O.RX, 0, 0, 0.3)).terms() == [("L", 1.0)]
    for two, one in ((O.RXX, O.RX), (O.RYY, O.RY), (O.RZZ, O.RZ)):
        for word in ("LZ", "LX", "LY", "ZL", "XL", "YL", "LL", "LI"):
            a = run(lossy(2, [word]), (two, 0, 1, 0.3))
            lost = word.index("L")
            b = lossy(2, [word])
            if word != "LL":
                run(b, (one, 1 - lost, 0, 0.3))
            assert a.terms() == b.terms(), (two, word)


def test_noise_skips_lost_qubits():
    # tests/noise.rs:180-194 and sum/noise.rs:17-131
    p3 = [0.01, 0.02, 0.03]
    assert run(lossy(1, ["L"]), (O.PAULI_ERROR, 0, 0, 0.0, p3)).terms() == [("L", 1.0)]
    assert run(lossy(1, ["L"]), (O.DEPOLARIZE1, 0, 0, 0.0, [0.1])).terms() == [("L", 1.0)]
    assert run(lossy(2, ["LX"]), (O.DEPOLARIZE2, 0, 1, 0.0, [0.1])).terms() == [("LX", 1.0)]
    p15 = [0.01 * (i + 1) for i in range(15)]
    assert run(lossy(2, ["XL"]), (O.PAULI_ERROR2, 0, 1, 0.0, p15)).terms() == [("XL", 1.0)]
    assert run(lossy(1, ["L"]), (O.AMPLITUDE_DAMPING, 0, 0, 0.0, [0.3])).terms() == [("L", 1.0)]


def _random_ops(rng, n, count, with_loss=True, avoid=()):
    kinds = [O.X, O.Y, O.Z, O.H, O.S, O.S_DAG, O.SQRT_X, O.SQRT_X_DAG, O.SQRT_Y, O.SQRT_Y_DAG, O.CNOT,
             O.CZ, O.CY, O.RX, O.RY, O.RZ, O.RXX, O.RYY, O.RZZ, O.PAULI_ERROR, O.DEPOLARIZE1,
             O.DEPOLARIZE2, O.PAULI_ERROR2, O.AMPLITUDE_DAMPING]
    if with_loss:
        kinds += [O.LOSS, O.LOSS, O.RESET_LOSS, O.RESET_LOSS, O.CORRELATED_LOSS, O.CORRELATED_LOSS]
    free = [q for q in range(n) if q not in avoid]
    p = O.Program()
    for _ in range(count):
        k = int(rng.choice(kinds))
        a, b = (int(v) for v in rng.choice(free, 2, replace=False))
        probs = list(rng.uniform(0.0, 0.06, 15))
        p.add(k, a, b, float(rng.uniform(-3, 3)), probs)
        if rng.random() < 0.5:
            p.truncate()
    return p


def _random_words(rng, n, count, lost=True):
    out = set()
    while len(out) < count:
        out.add("".join(rng.choice(list("IXYZL" if lost else "IXYZ"), p=[0.4, 0.15, 0.15, 0.15, 0.15]
                                   if lost else [0.4, 0.2, 0.2, 0.2]) for _ in range(n)))
    return sorted(out)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lossy_equals_plain_without_loss(seed):
    rng = np.random.default_rng(seed)
    n = 9
    words = _random_words(rng, n, 12, lost=False)
    coefs = list(rng.uniform(-1, 1, len(words)))
    prog = _random_ops(rng, n, 60, with_loss=False).array()
    a = LossyOracleSum(n, 1e-4, 6, None)
    a.add_strings(words, coefs)
    b = OracleSum(n, 1e-4, 6)
    b.add_strings(words, coefs)
    a.apply(prog)
    b.apply(prog)
    assert a.terms() == b.terms()
    assert a.overlap_with_zero() == b.overlap_with_zero()


@pytest.mark.parametrize("seed", [3, 4])
def test_kept_qubits_act_as_plain_planes(seed):
    # lost qubits 0 and 1 are never touched: the x/z planes evolve as a plain sum
    rng = np.random.default_rng(seed)
    n = 8
    words = _random_words(rng, n - 2, 10, lost=False)
    coefs = list(rng.uniform(-1, 1, len(words)))
    prog = _random_ops(rng, n, 50, with_loss=False, avoid=(0, 1)).array()
    a = LossyOracleSum(n, 0.0, None, None)
    a.add_strings(["LL" + w for w in words], coefs)
    b = OracleSum(n, 0.0, None)
    b.add_strings(["II" + w for w in words], coefs)
    a.apply(prog)
    b.apply(prog)
    assert [("LL" + w[2:], c) for w, c in b.terms()] == a.terms()


def test_loss_channels_scale_loss_free_sums():
    rng = np.random.default_rng(5)
    words = _random_words(rng, 5, 10, lost=False)
    coefs = list(rng.uniform(-1, 1, len(words)))
    p = 0.123
    s = lossy(5, dict(zip(words, coefs)))
    run(s, (O.LOSS, 2, 0, 0.0, [p]))
    assert s.terms() == [(w, c * (1.0 - p)) for w, c in zip(words, coefs)]
    pc = [0.05, 0.07, 0.11]
    s = lossy(5, dict(zip(words, coefs)))
    run(s, (O.CORRELATED_LOSS, 1, 3, 0.0, pc))
    assert s.terms() == [(w, c * (1.0 - pc[1] * 2.0 - pc[0])) for w, c in zip(words, coefs)]


def test_patterns_never_match_lost_qubits():
    s = lossy(3, {"ZIZ": 1.0, "ZLZ": 2.0, "LII": 4.0, "III": 8.0})
    assert s.overlap_with_zero() == 9.0
    assert s.trace("Z?*") == 9.0
    assert s.trace("Z0") == 0.0
    assert s.trace("Z0Z2") == 1.0


def test_preserve_restores_at_end():
    s = LossyOracleSum(2, 0.5, None, None)
    s.add_strings(["LZ", "XX", "ZL"], [0.1, 1.0, 0.2])
    s.set_preserve(["LZ"])
    run(s, (O.TRUNCATE,))
    assert s.terms() == [("XX", 1.0), ("LZ", 0.1)]
