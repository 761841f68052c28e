"""Pins the CPU oracle (tests/ CPU suite).

Each check is chosen so a plausible mistake (wrong sign, swapped x/z, wrong
qubit, dropped merge term, truncation in the wrong place) fails one of them:
  * golden constants of the reference's own tests
    (crates/ppvm-pauli-sum/tests/trotter.rs:104, tests/cnot.rs:27, tests/ghz.rs);
  * every Clifford / rotation on every 1- and 2-qubit Pauli against brute-force
    matrix conjugation U^dag P U with numpy/scipy;
  * random small circuits against a numpy state-vector expectation value;
  * the reference's exact surviving sets for truncation
    (tests/strategy.rs, tests/truncation_semantics.rs);
  * invariants: Frobenius norm conservation, Clifford |c| multiset.
"""

import itertools
import math

import numpy as np
import pytest
import scipy.linalg

from oracle.pyoracle import OracleSum
from ppvm_amd import ops as O

PAULI = {
    "I": np.eye(2, dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
}


def pmat(s: str) -> np.ndarray:
    m = np.array([[1.0 + 0j]])
    for ch in s:  # character 0 = qubit 0 = leftmost tensor factor
        m = np.kron(m, PAULI[ch])
    return m


def decompose(M: np.ndarray, n: int) -> dict:
    out = {}
    for t in itertools.product("IXYZ", repeat=n):
        s = "".join(t)
        c = np.trace(pmat(s) @ M) / 2**n
        if abs(c) > 1e-12:
            assert abs(c.imag) < 1e-12
            out[s] = c.real
    return out


def expm_pauli(G: str, theta: float) -> np.ndarray:
    return scipy.linalg.expm(-0.5j * theta * pmat(G))


P0 = np.array([[1, 0], [0, 0]], dtype=complex)
P1 = np.array([[0, 0], [0, 1]], dtype=complex)
GATE_1Q = {
    O.X: PAULI["X"], O.Y: PAULI["Y"], O.Z: PAULI["Z"],
    O.H: np.array([[1, 1], [1, -1]], dtype=complex) / math.sqrt(2),
    O.S: np.diag([1, 1j]), O.S_DAG: np.diag([1, -1j]),
    O.SQRT_X: expm_pauli("X", math.pi / 2), O.SQRT_X_DAG: expm_pauli("X", -math.pi / 2),
    O.SQRT_Y: expm_pauli("Y", math.pi / 2), O.SQRT_Y_DAG: expm_pauli("Y", -math.pi / 2),
}
GATE_2Q = {
    O.CNOT: np.kron(P0, PAULI["I"]) + np.kron(P1, PAULI["X"]),
    O.CZ: np.diag([1, 1, 1, -1]).astype(complex),
    O.CY: np.kron(P0, PAULI["I"]) + np.kron(P1, PAULI["Y"]),
}
ROT_G = {O.RX: "X", O.RY: "Y", O.RZ: "Z", O.RXX: "XX", O.RYY: "YY", O.RZZ: "ZZ"}


def oracle_single(n, s, prog_rows, tau=0.0):
    o = OracleSum(n, tau, None)
    o.add_strings([s], [1.0])
    p = O.Program()
    for row in prog_rows:
        p.add(*row)
    o.apply(p.array())
    return dict(o.terms())


# ---- golden constants -----------------------------------------------------------
def _trotter(tau):
    N, STEPS, TX, TZZ, NOISE = 4, 10, 0.1, 0.0125, [2.5e-5] * 3
    s = OracleSum(N, tau, None)
    s.add_strings(["".join("Z" if i == j else "I" for i in range(N)) for j in range(N)])
    p = O.Program()
    for _ in range(STEPS):
        for i in range(N):
            p.add(O.PAULI_ERROR, i, p=NOISE).truncate()
            p.add(O.RX, i, theta=TX).truncate()
        for i in range(N - 1):
            p.add(O.PAULI_ERROR, i + 1, p=NOISE).truncate()
            p.add(O.PAULI_ERROR, i, p=NOISE).truncate()
            p.add(O.RZZ, i, i + 1, theta=TZZ).truncate()
    s.apply(p.array())
    return s.overlap_with_zero()


def test_trotter_golden_and_truncation_fidelity():
    exact = _trotter(0.0)
    assert abs(exact - 2.1610566562692544) < 1e-9
    approx = _trotter(1e-6)
    assert abs(exact - approx) < 1e-6
    assert abs(exact - approx) > 1e-9  # truncation really dropped something


def test_cnot_golden():
    s = OracleSum(2, 0.0, None)
    s.add_strings(["ZZ"])
    p = O.Program()
    for q in (0, 1):
        p.add(O.RZ, q, theta=1.1).add(O.RY, q, theta=2.1).add(O.RZ, q, theta=1.1)
    p.add(O.CNOT, 0, 1).add(O.RX, 0, theta=2.1).add(O.RX, 1, theta=2.1)
    s.apply(p.array())
    v = s.overlap_with_zero()
    assert abs(v - 0.18803675917759355) <= 4 * np.finfo(float).eps * abs(v)


def test_ghz():
    assert oracle_single(1, "Y", [(O.H, 0)]) == {"Y": -1.0}
    s = OracleSum(2, 0.0, None)
    s.add_strings(["ZZ"])
    p = O.Program().add(O.CNOT, 0, 1).add(O.H, 0)
    s.apply(p.array())
    assert s.overlap_with_zero() == 1.0
    f = OracleSum(2, 0.0, None)
    f.add_strings(["ZI", "IZ", "ZZ", "II"])
    # forward GHZ preparation acting on the |00><00| expansion: conjugation by
    # the inverse circuit, i.e. cnot then h applied in Heisenberg order reversed
    f.apply(O.Program().add(O.H, 0).add(O.CNOT, 0, 1).array())
    assert dict(f.terms()) == {"XX": 1.0, "YY": -1.0, "ZZ": 1.0, "II": 1.0}


# ---- brute-force gate tables ----------------------------------------------------------
@pytest.mark.parametrize("code", list(GATE_1Q))
def test_clifford_1q_matches_conjugation(code):
    U = GATE_1Q[code]
    for s in "IXYZ":
        want = decompose(U.conj().T @ pmat(s) @ U, 1)
        got = oracle_single(1, s, [(code, 0)])
        assert got.keys() == want.keys() and all(abs(got[k] - want[k]) < 1e-12 for k in want), (s, got, want)


@pytest.mark.parametrize("code", list(GATE_2Q))
@pytest.mark.parametrize("order", [(0, 1), (1, 0)])
def test_clifford_2q_matches_conjugation(code, order):
    U = GATE_2Q[code]
    if order == (1, 0):  # control on qubit 1: swap tensor factors
        SW = np.eye(4)[[0, 2, 1, 3]]
        U = SW @ U @ SW
    for t in itertools.product("IXYZ", repeat=2):
        s = "".join(t)
        want = decompose(U.conj().T @ pmat(s) @ U, 2)
        got = oracle_single(2, s, [(code, order[0], order[1])])
        assert got.keys() == want.keys() and all(abs(got[k] - want[k]) < 1e-12 for k in want), (s, got, want)


@pytest.mark.parametrize("code", list(ROT_G))
@pytest.mark.parametrize("theta", [2.1, -0.7, math.pi / 2])
def test_rotations_match_conjugation(code, theta):
    G = ROT_G[code]
    n = len(G)
    U = expm_pauli(G, theta)
    for t in itertools.product("IXYZ", repeat=n):
        s = "".join(t)
        want = decompose(U.conj().T @ pmat(s) @ U, n)
        row = (code, 0, 1 if n == 2 else 0, theta)
        got = {k: v for k, v in oracle_single(n, s, [row]).items() if abs(v) > 1e-12}
        assert got.keys() == want.keys(), (s, got, want)
        assert all(abs(got[k] - want[k]) < 1e-12 for k in want), (s, got, want)


def test_rotation_explicit_reference_values():
    # crates/ppvm-pauli-sum/src/sum/rot1.rs:181-217 and sum/rot2.rs:304-327
    th = 2.1
    assert oracle_single(1, "Z", [(O.RX, 0, 0, th)]) == {"Z": math.cos(th), "Y": math.sin(th)}
    assert oracle_single(1, "Y", [(O.RX, 0, 0, th)]) == {"Y": math.cos(th), "Z": -math.sin(th)}
    assert oracle_single(1, "X", [(O.RY, 0, 0, th)]) == {"X": math.cos(th), "Z": math.sin(th)}
    t = 0.7
    assert oracle_single(2, "XI", [(O.RZZ, 0, 1, t)]) == {"XI": math.cos(t), "YZ": -math.sin(t)}
    assert oracle_single(2, "YI", [(O.RZZ, 0, 1, t)]) == {"YI": math.cos(t), "XZ": math.sin(t)}
    assert oracle_single(2, "ZZ", [(O.RZZ, 0, 1, t)]) == {"ZZ": 1.0}


def test_rotation_non_adjacent_addressing():
    got = oracle_single(5, "IXIII", [(O.RZZ, 1, 3, 0.4)])
    assert got == {"IXIII": math.cos(0.4), "IYIZI": -math.sin(0.4)}


# ---- random circuits vs state vector -------------------------------------------------------
def _statevector_expect(n, fwd_rows, obs_terms):
    psi = np.zeros(2**n, dtype=complex)
    psi[0] = 1.0

    def full(U, qs):
        # embed a 1- or 2-qubit gate acting on qubits qs (qubit 0 leftmost)
        if len(qs) == 1:
            mats = [U if q == qs[0] else np.eye(2) for q in range(n)]
            M = mats[0]
            for m in mats[1:]:
                M = np.kron(M, m)
            return M
        a, b = qs
        dim = 2**n
        M = np.zeros((dim, dim), dtype=complex)
        for col in range(dim):
            bits = [(col >> (n - 1 - q)) & 1 for q in range(n)]
            idx_in = bits[a] * 2 + bits[b]
            for idx_out in range(4):
                amp = U[idx_out, idx_in]
                if amp == 0:
                    continue
                nb = list(bits)
                nb[a], nb[b] = idx_out >> 1, idx_out & 1
                row = sum(v << (n - 1 - q) for q, v in enumerate(nb))
                M[row, col] += amp
        return M

    for code, a, b, th in fwd_rows:
        if code in GATE_1Q:
            psi = full(GATE_1Q[code], [a]) @ psi
        elif code in GATE_2Q:
            psi = full(GATE_2Q[code], [a, b]) @ psi
        else:
            G = ROT_G[code]
            U = expm_pauli(G, th)
            psi = full(U, [a] if len(G) == 1 else [a, b]) @ psi
    Om = sum(c * pmat(s) for s, c in obs_terms)
    return (psi.conj() @ Om @ psi).real


@pytest.mark.parametrize("seed", range(6))
def test_random_circuit_matches_statevector(seed):
    rng = np.random.default_rng(seed)
    n = 4
    codes = list(GATE_1Q) + list(GATE_2Q) + list(ROT_G)
    rows = []
    for _ in range(25):
        c = codes[rng.integers(len(codes))]
        a = int(rng.integers(n))
        b = int((a + 1 + rng.integers(n - 1)) % n)
        two = c in GATE_2Q or c in (O.RXX, O.RYY, O.RZZ)
        rows.append((c, a, b if two else 0, float(rng.uniform(-3, 3))))
    obs = [("".join("IXYZ"[v] for v in rng.integers(0, 4, n)), float(rng.normal())) for _ in range(3)]
    obs = list(dict(obs).items())
    s = OracleSum(n, 0.0, None)
    s.add_strings([o for o, _ in obs], [c for _, c in obs])
    p = O.Program()
    for r in reversed(rows):
        p.add(*r)
    s.apply(p.array())
    assert abs(s.overlap_with_zero() - _statevector_expect(n, rows, obs)) < 1e-10


# ---- truncation semantics --------------------------------------------------------------------
def _sum(n, terms, tau=0.0, maxw=None):
    s = OracleSum(n, tau, maxw)
    s.add_strings([t for t, _ in terms], [c for _, c in terms])
    return s


def test_strategy_surviving_sets():
    # tests/strategy.rs:10-91
    s = _sum(2, [("ZZ", 1.0), ("XX", 1e-3), ("YY", 1e-5), ("XZ", -0.5)], tau=1e-4)
    s.apply(O.Program().truncate().array())
    assert dict(s.terms()) == {"ZZ": 1.0, "XX": 1e-3, "XZ": -0.5}
    s = _sum(3, [("XXX", 1.0), ("ZZI", 1e-4), ("YZY", -10.0), ("IXX", -0.1)], maxw=2)
    s.apply(O.Program().truncate().array())
    assert dict(s.terms()) == {"ZZI": 1e-4, "IXX": -0.1}
    s = _sum(3, [("ZZI", 1.0), ("XIX", 1e-3), ("IYY", 1e-5), ("XIZ", -0.5), ("XXX", 1.0),
                 ("ZIZ", 1e-4), ("YZY", -10.0), ("IXX", -0.1)], tau=1e-4, maxw=2)
    s.apply(O.Program().truncate().array())
    assert dict(s.terms()) == {"ZZI": 1.0, "XIX": 1e-3, "XIZ": -0.5, "ZIZ": 1e-4, "IXX": -0.1}
    # insertion order is preserved by the truncation
    assert [k for k, _ in s.terms()] == ["ZZI", "XIX", "XIZ", "ZIZ", "IXX"]


def test_threshold_is_strict():
    s = _sum(1, [("Z", 1e-4), ("X", 0.99e-4)], tau=1e-4)
    s.apply(O.Program().truncate().array())
    assert dict(s.terms()) == {"Z": 1e-4}


def test_deferred_truncation_semantics():
    # tests/truncation_semantics.rs:35-79
    TAU, TH = 0.05, 0.03
    a = _sum(1, [("Z", 1.0)], tau=TAU)
    a.apply(O.Program().add(O.RX, 0, theta=TH).add(O.RX, 0, theta=TH).truncate().array())
    e = _sum(1, [("Z", 1.0)])
    e.apply(O.Program().add(O.RX, 0, theta=TH).add(O.RX, 0, theta=TH).array())
    ref = {k: v for k, v in e.terms() if abs(v) >= TAU}
    assert "Y" in ref
    assert dict(a.terms()) == ref


def test_preserve_strings_restored_at_end():
    s = OracleSum(2, 1e-3, None)
    s.set_preserve(["XX"])
    s.add_strings(["XX", "ZZ", "YY"], [1e-5, 1.0, 1e-6])
    s.apply(O.Program().truncate().array())
    assert s.terms() == [("ZZ", 1.0), ("XX", 1e-5)]


# ---- invariants ----------------------------------------------------------------------------------
def test_norm_conserved_without_truncation():
    rng = np.random.default_rng(3)
    n = 12
    s = OracleSum(n, 0.0, None)
    s.add_strings(["".join("IXYZ"[v] for v in rng.integers(0, 4, n)) for _ in range(5)], rng.normal(size=5))
    x0 = np.sum(s.export()[2] ** 2)
    p = O.Program()
    for _ in range(40):
        a, b = rng.choice(n, 2, replace=False)
        p.add([O.RX, O.RY, O.RZ, O.RZZ, O.RXX, O.RYY, O.CNOT, O.H][rng.integers(8)], int(a), int(b),
              theta=float(rng.uniform(-2, 2)))
    s.apply(p.array())
    assert abs(np.sum(s.export()[2] ** 2) - x0) < 1e-9 * x0


def test_clifford_preserves_count_and_magnitudes():
    rng = np.random.default_rng(4)
    n = 10
    s = OracleSum(n, 0.0, None)
    strs = list({"".join("IXYZ"[v] for v in rng.integers(0, 4, n)) for _ in range(50)})
    s.add_strings(strs, rng.normal(size=len(strs)))
    mags = np.sort(np.abs(s.export()[2]))
    p = O.Program()
    for _ in range(60):
        a, b = rng.choice(n, 2, replace=False)
        p.add([O.H, O.S, O.CNOT, O.CZ, O.CY, O.SQRT_X, O.SQRT_Y_DAG][rng.integers(7)], int(a), int(b))
    s.apply(p.array())
    assert len(s) == len(strs)
    assert np.array_equal(np.sort(np.abs(s.export()[2])), mags)


def test_sharded_variant_matches_single_thread():
    from ppvm_amd import workloads as WL
    w = WL.truncate_steps(WL.lattice_tfim(), 4)
    a = _sum(w.n_qubits, w.observable, w.min_abs_coeff, w.max_pauli_weight)
    b = _sum(w.n_qubits, w.observable, w.min_abs_coeff, w.max_pauli_weight)
    a.apply(w.program)
    b.apply_sharded(w.program, 4)
    assert dict(a.terms()) == dict(b.terms())
    assert a.term_updates() == b.term_updates()


# ---- overlap between two sums ------------------------------------------------------------------
def test_overlap_known_answers():
    """crates/ppvm-pauli-sum/src/sum/trace.rs:52-80."""
    a = _sum(4, [("IIII", 1.0), ("XIII", 2.0)])
    empty = _sum(4, [])
    assert a.overlap(empty) == 0.0 and empty.overlap(a) == 0.0
    s = _sum(4, [("XIII", 3.0)])
    assert s.overlap(s) == 9.0
    assert _sum(4, [("XIII", 1.0)]).overlap(_sum(4, [("YIII", 1.0)])) == 0.0
    b = _sum(4, [("IIII", 3.0), ("XIII", 4.0)])
    assert a.overlap(b) == 1.0 * 3.0 + 2.0 * 4.0


def test_overlap_is_hilbert_schmidt_product():
    """overlap(A, B) = Tr(A B) / 2^n for real Pauli sums, after random circuits."""
    rng = np.random.default_rng(12)
    n = 4
    for trial in range(4):
        sums = []
        for side in range(2):
            strs = list({"".join("IXYZ"[v] for v in rng.integers(0, 4, n)) for _ in range(12)})
            s = _sum(n, list(zip(strs, rng.normal(size=len(strs)).tolist())))
            p = O.Program()
            for _ in range(10):
                a, b = rng.choice(n, 2, replace=False)
                code = [O.H, O.CNOT, O.RX, O.RZZ, O.RY][rng.integers(5)]
                p.add(code, int(a), int(b), theta=float(rng.uniform(-2, 2)))
            s.apply(p.array())
            sums.append(s)
        A = sum(c * pmat(t) for t, c in sums[0].terms())
        B = sum(c * pmat(t) for t, c in sums[1].terms())
        hs = np.trace(A @ B).real / 2**n
        assert abs(sums[0].overlap(sums[1]) - hs) < 1e-12
        assert abs(sums[1].overlap(sums[0]) - hs) < 1e-12


# ---- trace(pattern) ----------------------------------------------------------------------------
def _contains(pattern, word):
    return OracleSum(len(word), 0.0, None).pattern_contains(pattern, [word])[0]


def test_pattern_reference_vectors():
    """crates/ppvm-pauli-word/src/pattern/contains.rs:165-213 and parse.rs docs."""
    assert _contains("X0Y1Z2", "XYZ")
    assert _contains("X?0Y1Z2", "XYZ") and _contains("X?0Y1Z2", "IYZ")
    assert _contains("[XY]0Y1Z2", "XYZ") and _contains("[XY]0Y1Z2", "YYZ")
    for w in ("XYZ", "YYZ", "IYZ"):
        assert _contains("[XY]?0Y1Z2", w)
    for w in ("XYX", "YYX", "IYX"):
        assert _contains("[XY]?*", w)
    for w in ("XYZ", "YYZ", "IYZ"):
        assert _contains("[XY]?{2}Z2", w)
    assert not _contains("Z?*", "XYY")


def test_pattern_semantics_edges():
    # anchored positions need identities before them and a match at the anchor
    assert _contains("Z3", "IIIZ") and not _contains("Z3", "IXIZ") and not _contains("Z3", "IIIX")
    assert not _contains("Z5", "IIIZ")              # anchor past the word: no match
    assert _contains("Z1Z3", "IZIZ") and not _contains("Z3Z1", "IZIZ")  # anchors consume in order
    # trailing symbols must be identities
    assert not _contains("X0", "XZ") and _contains("X0", "XI")
    # empty pattern: the identity word only
    assert _contains("", "III") and not _contains("", "IXI")
    # the "any Pauli or identity" forms accept the identity only (contains.rs:39)
    assert _contains("_*", "III") and not _contains("_*", "IXI")
    assert _contains("[XYZ]?*", "III") and not _contains("[XYZ]?*", "ZII")
    assert _contains("[XYZ]*", "XYZ") and not _contains("[XYZ]*", "XIZ")
    # counted repetition: exactly `count` matches consumed, the rest must be identity
    assert _contains("X{2}", "XXI") and not _contains("X{2}", "XII") and not _contains("X{2}", "XXX")
    # count 0 (contains.rs:94-116): matches are consumed past the count; a later
    # non-match ends the item successfully, running out of symbols does not
    assert _contains("X{0}", "III") and _contains("X{0}", "XII") and not _contains("X{0}", "XXX")
    # whitespace at the ends is trimmed
    assert _contains("  Z?*  ", "ZIZ")


@pytest.mark.parametrize("bad", ["A", "[]", "X", "[XW]0", "X{a}", "X*Y", "X 0", "[XY"])
def test_pattern_parse_errors(bad):
    with pytest.raises(ValueError):
        OracleSum(4, 0.0, None).trace(bad)


def test_trace_closed_forms():
    rng = np.random.default_rng(21)
    n = 6
    strs = list({"".join("IXYZ"[v] for v in rng.integers(0, 4, n)) for _ in range(300)}) + ["I" * n]
    coefs = rng.normal(size=len(strs))
    s = _sum(n, list(zip(strs, coefs.tolist())))
    d = dict(zip(strs, coefs.tolist()))
    assert s.trace("Z?*") == s.overlap_with_zero()
    assert s.trace("") == d["I" * n]
    assert s.trace("_*") == d["I" * n]
    for w in strs[:40]:  # a fully anchored pattern selects exactly one word
        assert s.trace("".join(f"{ch}{q}" for q, ch in enumerate(w) if ch != "I")) == d[w]
    # X?* sums the words made of I and X only, in insertion order
    want = 0.0
    for w, c in zip(strs, coefs.tolist()):
        if set(w) <= {"I", "X"}:
            want += c
    assert s.trace("X?*") == want


# ---- amplitude damping -----------------------------------------------------------------------
def _damp(s, q, gamma):
    s.apply(O.Program().add(O.AMPLITUDE_DAMPING, q, p=[gamma]).array())


def test_amplitude_damping_reference_vectors():
    """ppvm-python/test/test_basics.py:390-414 (restated; the Python class
    truncates after the channel with its default min_abs_coeff 1e-10)."""
    def run(word, q, g):
        s = _sum(2, [(word, 1.0)], tau=1e-10)
        s.apply(O.Program().add(O.AMPLITUDE_DAMPING, q, p=[g]).truncate().array())
        return dict(s.terms())
    assert run("ZI", 0, 0.0) == {"ZI": 1.0}
    g = 0.2
    t = run("ZI", 0, g)
    assert t["ZI"] == pytest.approx(1.0 - g) and t["II"] == pytest.approx(g)
    assert run("XI", 0, g) == {"XI": pytest.approx((1.0 - g) ** 0.5)}
    assert run("ZI", 1, g) == {"ZI": 1.0}


def test_amplitude_damping_matches_kraus_adjoint():
    """E^dag(P) = sum_i K_i^dag P K_i with K0 = diag(1, sqrt(1-g)), K1 = sqrt(g)|0><1|,
    on random 3-qubit sums (I and Z versions of a word present, so children merge)."""
    rng = np.random.default_rng(33)
    n = 3
    for trial in range(6):
        g = float(rng.uniform(0.05, 0.95))
        q = int(rng.integers(0, n))
        strs = list({"".join("IXYZ"[v] for v in rng.integers(0, 4, n)) for _ in range(14)})
        coefs = rng.normal(size=len(strs))
        s = _sum(n, list(zip(strs, coefs.tolist())))
        _damp(s, q, g)
        K0 = np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=complex)
        K1 = np.array([[0, math.sqrt(g)], [0, 0]], dtype=complex)

        def on_q(K):
            m = np.array([[1.0 + 0j]])
            for i in range(n):
                m = np.kron(m, K if i == q else np.eye(2))
            return m
        A = sum(c * pmat(w) for w, c in zip(strs, coefs))
        E = on_q(K0).conj().T @ A @ on_q(K0) + on_q(K1).conj().T @ A @ on_q(K1)
        want = decompose(E, n)
        got = dict(s.terms())
        assert set(k for k, v in got.items() if abs(v) > 1e-12) == set(want)
        for k, v in want.items():
            assert abs(got[k] - v) < 1e-12
