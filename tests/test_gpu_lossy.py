"""Lossy sums on the GPU (ppvm_lossy_* through the C ABI) against the CPU
oracle: identical words, order and coefficient bits after every segment of
random circuits that mix Cliffords, rotations, noise, loss, correlated loss,
reset-loss and truncation; the reference's known answers through the Python
``LossyPauliSum`` (ppvm-python/test/test_loss.py, crates/ppvm-pauli-sum/tests/loss.rs);
preserve sets, loss-weight caps, patterns, overlaps and error paths.
"""

import numpy as np
import pytest

from oracle.pyoracle import LossyOracleSum
from ppvm_amd import ops as O

pytestmark = pytest.mark.gpu

KINDS = [O.X, O.Y, O.Z, O.H, O.S, O.S_DAG, O.SQRT_X, O.SQRT_X_DAG, O.SQRT_Y, O.SQRT_Y_DAG, O.CNOT, O.CZ,
         O.CY, O.RX, O.RY, O.RZ, O.RXX, O.RYY, O.RZZ, O.PAULI_ERROR, O.DEPOLARIZE1, O.DEPOLARIZE2,
         O.PAULI_ERROR2, O.AMPLITUDE_DAMPING, O.LOSS, O.LOSS, O.RESET_LOSS, O.RESET_LOSS,
         O.CORRELATED_LOSS, O.CORRELATED_LOSS]


def _handle(n, tau=0.0, maxw=None, maxlw=None):
    from ppvm_amd._native import LossyHandle
    return LossyHandle(n, tau, maxw, maxlw)


def _words(rng, n, count, support):
    out = set()
    while len(out) < count:
        chars = ["I"] * n
        for q in rng.choice(support, min(len(support), int(rng.integers(1, 5))), replace=False):
            chars[int(q)] = "IXYZL"[int(rng.integers(0, 5))]
        out.add("".join(chars))
    return sorted(out)


def _program(rng, support, count, trunc_p=0.5):
    p = O.Program()
    for _ in range(count):
        k = int(rng.choice(KINDS))
        a, b = (int(v) for v in rng.choice(support, 2, replace=False))
        p.add(k, a, b, float(rng.uniform(-3, 3)), list(rng.uniform(0.0, 0.08, 15)))
        if rng.random() < trunc_p:
            p.truncate()
    return p.array()


def _same(g, o, ctx):
    gx, gz, gl, gc = g.export()
    ox, oz, ol, oc = o.export()
    assert len(gc) == len(oc), f"{ctx}: {len(gc)} GPU terms vs {len(oc)} oracle terms"
    assert np.array_equal(gx, ox) and np.array_equal(gz, oz) and np.array_equal(gl, ol), \
        f"{ctx}: words or order differ"
    bad = np.nonzero(gc.view(np.uint64) != oc.view(np.uint64))[0]
    assert len(bad) == 0, f"{ctx}: {len(bad)} coefficients differ, first {gc[bad[0]]!r} vs {oc[bad[0]]!r}"


def _load(g, o, words, coefs):
    from ppvm_amd.ops import pack_lossy_strings
    x, z, l = pack_lossy_strings(words, o.n_qubits)
    g.add_words(x, z, l, np.asarray(coefs))
    o.add_lossy_words(x, z, l, np.asarray(coefs))


@pytest.mark.parametrize("n,seed,tau,maxw,maxlw", [
    (6, 0, 0.0, None, None), (9, 1, 1e-4, None, 2), (12, 2, 1e-6, 7, None),
    (70, 3, 1e-5, 8, 3), (130, 4, 1e-5, None, 2), (300, 5, 1e-5, 10, 2)])
def test_random_lossy_circuits(n, seed, tau, maxw, maxlw):
    rng = np.random.default_rng(seed)
    support = np.array(sorted(rng.choice(n, min(n, 8), replace=False)))
    print(f"lossy n={n} seed={seed}", flush=True)
    g, o = _handle(n, tau, maxw, maxlw), LossyOracleSum(n, tau, maxw, maxlw)
    words = _words(rng, n, 24, support)
    _load(g, o, words, rng.uniform(-1, 1, len(words)))
    prog = _program(rng, support, 160)
    for s in range(0, len(prog), 20):
        seg = prog[s:s + 20]
        g.apply(seg)
        o.apply(seg)
        _same(g, o, f"n={n} seed={seed} ops<{s + 20}")
        assert abs(g.overlap_with_zero() - o.overlap_with_zero()) <= 1e-10
    for pat in ("Z?*", "[XZ]?*", "Z0", "X?{3}Z?*"):
        assert abs(g.trace(pat) - o.trace(pat)) <= 1e-10, pat


def test_many_terms_grow_and_merge():
    # tens of thousands of words: multi-block scans, arena growth during a
    # branch op, and every reset / correlated child landing on existing words
    rng = np.random.default_rng(7)
    n = 16
    support = np.arange(n)
    g, o = _handle(n, 1e-9), LossyOracleSum(n, 1e-9)
    words = sorted({"".join(rng.choice(list("IXYZL"), n, p=[0.3, 0.15, 0.15, 0.25, 0.15])) for _ in range(8000)})
    _load(g, o, words, rng.uniform(-1, 1, len(words)))
    prog = O.Program()
    for q in range(0, 4):
        prog.add(O.RESET_LOSS, q).add(O.CORRELATED_LOSS, q, q + 8, p=[0.02, 0.03, 0.05]).add(O.LOSS, q + 8, p=[0.1])
        prog.add(O.RZZ, q, q + 1, theta=0.3).add(O.RX, q + 8, theta=0.2).truncate()
    arr = prog.array()
    for s in range(0, len(arr), 6):
        g.apply(arr[s:s + 6])
        o.apply(arr[s:s + 6])
        _same(g, o, f"grow ops<{s + 6}")
    assert len(o) > 2 * len(words)


def test_preserve_and_overlap_and_clone():
    n = 10
    rng = np.random.default_rng(11)
    support = np.arange(n)
    from ppvm_amd.ops import pack_lossy_strings
    keep = ["LZIIIIIIII", "XXIIIIIIIL"]
    a, oa = _handle(n, 0.05), LossyOracleSum(n, 0.05)
    for h in (a,):
        h.set_preserve(*pack_lossy_strings(keep, n))
    oa.set_preserve(keep)
    words = _words(rng, n, 40, support) + keep
    coefs = list(rng.uniform(-1, 1, len(words) - 2)) + [0.01, -0.02]
    _load(a, oa, words, coefs)
    prog = _program(rng, support, 60, trunc_p=0.7)
    a.apply(prog)
    oa.apply(prog)
    _same(a, oa, "preserve")
    b, ob = a.clone(), oa.copy()
    prog2 = _program(np.random.default_rng(12), support, 30)
    b.apply(prog2)
    ob.apply(prog2)
    _same(b, ob, "clone")
    _same(a, oa, "original after clone")
    assert abs(a.overlap(b) - oa.overlap(ob)) <= 1e-10


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ops_that_change_nothing_are_skipped_exactly(seed):
    """The engine skips an op whose rule table leaves every local state the
    sum may hold on its qubits unchanged, and a truncation with nothing
    changed since the last one.  Words live on qubits 0-5 (with lost qubits),
    and the program mixes ops on qubits no term touches, rotations on qubits
    holding only I and L, zero-probability noise, repeated truncations of a
    preserve set whose words fall below the threshold, and Cliffords that move
    states onto new qubits before rotations act there; the oracle runs every op."""
    n = 20
    rng = np.random.default_rng(500 + seed)
    from ppvm_amd.ops import pack_lossy_strings
    keep = ["ZLIIIIIIIIIIIIIIIIII", "IIXIIIIIIIIIIIIIIIII"]
    g, o = _handle(n, 0.02, 6, 2), LossyOracleSum(n, 0.02, 6, 2)
    g.set_preserve(*pack_lossy_strings(keep, n))
    o.set_preserve(keep)
    words = _words(rng, n, 30, np.arange(6)) + keep
    _load(g, o, words, list(rng.uniform(-1, 1, len(words) - 2)) + [0.01, -0.015])
    p = O.Program()
    for _ in range(120):
        r = rng.random()
        if r < 0.25:  # qubits no term touches yet (until a Clifford moves a state there)
            a, b = (int(v) for v in rng.choice(np.arange(10, 20), 2, replace=False))
        else:
            a, b = (int(v) for v in rng.choice(np.arange(0, 12), 2, replace=False))
        k = int(rng.choice([O.RX, O.RY, O.RZ, O.RZZ, O.RXX, O.CNOT, O.H, O.PAULI_ERROR, O.LOSS, O.RESET_LOSS]))
        probs = [0.0] * 15 if (k == O.PAULI_ERROR and rng.random() < 0.5) else list(rng.uniform(0.0, 0.05, 15))
        p.add(k, a, b, float(rng.uniform(-2, 2)), probs)
        for _ in range(int(rng.integers(0, 3))):
            p.truncate()
    prog = p.array()
    for s in range(0, len(prog), 15):
        g.apply(prog[s:s + 15])
        o.apply(prog[s:s + 15])
        _same(g, o, f"seed {seed} ops {s}")
    c, oc = g.clone(), o.copy()
    tail = O.Program()
    tail.truncate()
    tail.add(O.RZ, 15, 0, 0.4)
    tail.truncate()
    c.apply(tail.array())
    oc.apply(tail.array())
    _same(c, oc, "clone")


def test_python_lossy_known_answers():
    from ppvm_amd import LossyPauliSum
    p = 0.1
    ps = LossyPauliSum.new(n_qubits=2, terms=["ZZ"])
    ps.reset_loss_channel(0)
    ps.loss_channel(0, p)
    ps.cnot(0, 1)
    ps.h(0)
    assert abs(ps.overlap_with_zero() - (1 - p)) < 1e-9
    for obs, expect in (("ZZ", 1.0), ("XX", 1 - 0.1)):
        ps = LossyPauliSum.new(n_qubits=2, terms=[obs])
        ps.reset_loss_channel(0)
        ps.reset_loss_channel(1)
        ps.correlated_loss_channel(0, 1, [0.1, 0.0, 0.0])
        ps.cnot(0, 1)
        ps.h(0)
        assert abs(ps.overlap_with_zero() - expect) < 1e-9
    ps = LossyPauliSum.new(n_qubits=1, terms=[("I", 2.0), ("Z", 3.0)])
    ps.reset_loss_channel(0)
    assert ps.terms == [("I", 2.0), ("Z", 3.0), ("L", 5.0)]
    assert dict(ps.weights()) == {"I": 0, "Z": 1, "L": 1}
    ps = LossyPauliSum.new(n_qubits=3, terms=["ZZZ"], max_loss_weight=2)
    for q in range(3):
        ps.reset_loss_channel(q, truncate=False)
    for q in range(3):
        ps.loss_channel(q, 0.1, truncate=False)
    before = len(ps)
    ps.truncate()
    assert len(ps) == before - 1
    assert max(w for _, w in ps.loss_weights()) == 2


def test_errors():
    from ppvm_amd import LossyPauliSum, PauliSum
    from ppvm_amd.ops import pack_lossy_strings
    g = _handle(4)
    x, z, l = pack_lossy_strings(["XIII"], 4)
    with pytest.raises(ValueError):
        g.add_words(x, z, x, np.ones(1))  # X and L on one qubit
    with pytest.raises(ValueError):
        g.apply(O.Program().add(O.LOSS, 4, p=[0.1]).array())
    with pytest.raises(ValueError):
        g.apply(O.Program().add(O.CORRELATED_LOSS, 1, 1, p=[0.1, 0, 0]).array())
    with pytest.raises(ValueError):
        g.trace("Q0")
    plain = PauliSum.new(4, "ZIII")
    with pytest.raises(ValueError):
        plain.apply(O.Program().add(O.LOSS, 0, p=[0.1]).array())
    with pytest.raises(ValueError):
        LossyPauliSum.new(2, "ZQ")
    with pytest.raises(TypeError):
        LossyPauliSum.new(2, "ZZ").overlap(PauliSum.new(2, "ZZ"))


def test_empty_sum_and_export_capacity():
    import ctypes

    from ppvm_amd import _native
    g = _handle(5, 0.5)
    o = LossyOracleSum(5, 0.5)
    _load(g, o, ["ZLIII", "XIIII"], [0.1, 0.2])
    prog = O.Program().truncate().add(O.RESET_LOSS, 0).add(O.CORRELATED_LOSS, 1, 2, p=[0.1, 0.1, 0.1])
    prog.add(O.RXX, 0, 1, theta=0.4).add(O.LOSS, 3, p=[0.2]).truncate()
    g.apply(prog.array())
    o.apply(prog.array())
    assert len(g) == 0 and len(o) == 0
    assert g.overlap_with_zero() == 0.0 and g.trace("Z?*") == 0.0
    x, z, l, c = g.export()
    assert len(c) == 0
    # a buffer smaller than the sum: nothing written, the length reported, EINVAL
    h = _handle(3)
    _load(h, LossyOracleSum(3), ["ZLI", "IIZ"], [1.0, 2.0])
    buf = np.zeros((1, 1), dtype=np.uint64)
    cbuf = np.zeros(1)
    got = ctypes.c_size_t()
    rc = _native.lib().ppvm_lossy_export_terms(h.h, _native.u64p(buf), _native.u64p(buf), _native.u64p(buf),
                                                _native.f64p(cbuf), 1, ctypes.byref(got))
    assert rc == _native.PPVM_EINVAL and got.value == 2


@pytest.mark.parametrize("case", ["aux_smaller", "aux_larger", "single_ll", "random_lost"])
def test_correlated_loss_merge_order_and_association(case):
    """correlated_loss_channel goes through map_insert_multiple
    (crates/ppvm-pauli-sum/src/sum/data.rs:234-247): children are summed among
    themselves, added to an existing word once (p + S), and when the distinct
    children outnumber the sum's words the children's order comes first.  GPU
    words, order and coefficient bits equal the oracle's in every case."""
    rng = np.random.default_rng(7)
    pr = [0.3, 0.1, 0.7]
    if case == "aux_smaller":
        words, coefs = ["II", "IL", "LI", "LL"], [1e16, 4.0, 4.0, 1.3]
        n = 2
    elif case == "aux_larger":
        words, coefs = ["LL", "IL"], [2.0, 0.5]
        n = 2
    elif case == "single_ll":
        words, coefs = ["LL"], [1.0]
        n = 2
    else:
        # many words with both target qubits lost: three distinct children each
        n = 12
        words = _words(rng, n, 300, list(range(2, n)))
        words = sorted({w[:3] + "L" + w[4:6] + "L" + w[7:] for w in words})
        coefs = list(rng.uniform(-1, 1, len(words)))
    g, o = _handle(n), LossyOracleSum(n, 0.0, None, None)
    _load(g, o, words, coefs)
    a, b = (0, 1) if n == 2 else (3, 6)
    prog = O.Program().add(O.CORRELATED_LOSS, a, b, 0.0, pr).array()
    g.apply(prog)
    o.apply(prog)
    _same(g, o, case)
    # and once more on the result (the first order decides the second's input)
    g.apply(prog)
    o.apply(prog)
    _same(g, o, case + " twice")
