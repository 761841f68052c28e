"""Noise channels and the reference's golden circuits, on the oracle (CPU) and
on the HIP engine (GPU).

two_qubit_pauli_error (crates/ppvm-pauli-sum/src/sum/noise.rs:52-107) scales
a term P_a P_b by 1 - 2 sum_j p_j over the channel Paulis Q_j (in the order
IX, IY, IZ, XI, XX, XY, XZ, YI, YX, YY, YZ, ZI, ZX, ZY, ZZ) that anticommute
with P_a P_b.  That symplectic closed form pins every one of the 16 rows:
a dropped, extra or transposed index, or qubits a and b swapped, changes the
factor.  The known answers of crates/ppvm-pauli-sum/tests/noise.rs:8-216 are
restated on both sides, and the reference's golden expectation values
(tests/trotter.rs:104, tests/cnot.rs:27) are reproduced on the HIP path.
"""

import itertools
import math

import numpy as np
import pytest

from oracle.pyoracle import OracleSum
from ppvm_amd import ops as O

CHANNEL = ["IX", "IY", "IZ", "XI", "XX", "XY", "XZ", "YI", "YX", "YY", "YZ", "ZI", "ZX", "ZY", "ZZ"]


def anticommute(p: str, q: str) -> bool:
    """Two Pauli strings anticommute iff an odd number of positions hold two
    different non-identity letters."""
    return sum(a != "I" and b != "I" and a != b for a, b in zip(p, q)) % 2 == 1


def closed_form(word: str, p) -> float:
    return 1.0 - 2.0 * math.fsum(pj for q, pj in zip(CHANNEL, p) if anticommute(word, q))


def oracle_factor(word: str, p, n: int = 2, a: int = 0, b: int = 1) -> float:
    s = OracleSum(n, 0.0, None)
    full = ["I"] * n
    full[a], full[b] = word[0], word[1]
    s.add_strings(["".join(full)], [1.0])
    s.apply(O.Program().add(O.PAULI_ERROR2, a, b, p=p).array())
    (_, c), = s.terms()
    return c


WORDS = ["".join(t) for t in itertools.product("IXYZ", repeat=2)]


def test_anticommute_helper():
    # Z anticommutes with X and Y, commutes with I and Z; XX commutes with YY
    assert anticommute("Z", "X") and anticommute("Z", "Y")
    assert not anticommute("Z", "I") and not anticommute("Z", "Z")
    assert not anticommute("XX", "YY") and anticommute("XI", "ZZ")
    # every non-identity two-qubit Pauli anticommutes with exactly 8 of the 15
    for w in WORDS[1:]:
        assert sum(anticommute(w, q) for q in CHANNEL) == 8


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_two_qubit_pauli_error_rows_match_symplectic_closed_form(seed):
    rng = np.random.default_rng(seed)
    p = list(rng.uniform(0.0, 1.0 / 15.0, size=15))
    for w in WORDS:
        got = oracle_factor(w, p)
        want = closed_form(w, p)
        # same terms, another summation order: a few ulp apart at most
        assert abs(got - want) <= 8 * np.finfo(float).eps, (w, got, want)


def _factor_abs(word: str, p, a: int, b: int) -> float:
    """Factor of the channel on qubits (a, b) for a term given in absolute positions."""
    s = OracleSum(len(word), 0.0, None)
    s.add_strings([word], [1.0])
    s.apply(O.Program().add(O.PAULI_ERROR2, a, b, p=p).array())
    (_, c), = s.terms()
    return c


def test_two_qubit_pauli_error_qubit_order():
    # the channel's first letter acts on a, the second on b: IX on (0, 2) puts X
    # on qubit 2, IX on (2, 0) puts X on qubit 0
    p = [0.0] * 15
    p[0] = 0.3  # IX
    assert _factor_abs("ZII", p, 0, 2) == 1.0
    assert _factor_abs("ZII", p, 2, 0) == pytest.approx(0.4)
    assert _factor_abs("IIZ", p, 0, 2) == pytest.approx(0.4)
    assert _factor_abs("IIZ", p, 2, 0) == 1.0


# ---- crates/ppvm-pauli-sum/tests/noise.rs known answers ---------------------------------
def _p(**kv):
    p = [0.0] * 15
    for name, v in kv.items():
        p[CHANNEL.index(name)] = v
    return p


# (start word, channel probabilities, expected {word: coefficient}); the expected
# sides are the reference's own: x(1), y(1), z(1) Cliffords, rxx(pi) then
# truncate, unchanged, and multiplication by -1 (tests/noise.rs:8-94)
NOISE_RS_CASES = [
    ("IZ", _p(IX=1.0), {"IZ": -1.0}),
    ("IZ", _p(IY=1.0), {"IZ": -1.0}),
    ("IX", _p(IZ=1.0), {"IX": -1.0}),
    ("YI", _p(XX=1.0), {"YI": -1.0}),
    ("YZ", _p(XX=1.0), {"YZ": 1.0}),
    ("XI", _p(ZZ=1.0), {"XI": -1.0}),
    ("XI", _p(XX=1.0, ZZ=1.0), {"XI": -1.0}),
]


def _reference_side(start):
    """The right-hand sides as the reference builds them (state2)."""
    out = {}
    for start_w, op in [("IZ", (O.X, 1)), ("IX", (O.Z, 1))]:
        s = OracleSum(2, 1e-12, None)
        s.add_strings([start_w], [1.0])
        s.apply(O.Program().add(op[0], op[1]).array())
        out[start_w] = dict(s.terms())
    s = OracleSum(2, 1e-12, None)
    s.add_strings(["YI"], [1.0])
    s.apply(O.Program().add(O.RXX, 0, 1, theta=math.pi).truncate().array())
    out["YI"] = dict(s.terms())
    return out[start]


def test_noise_rs_reference_sides_agree_with_expectations():
    assert _reference_side("IZ") == {"IZ": -1.0}   # x(1) on IZ
    assert _reference_side("IX") == {"IX": -1.0}   # z(1) on IX
    assert _reference_side("YI") == {"YI": -1.0}   # rxx(pi) then truncate


@pytest.mark.parametrize("start,p,want", NOISE_RS_CASES)
def test_noise_rs_on_oracle(start, p, want):
    s = OracleSum(2, 1e-12, None)
    s.add_strings([start], [1.0])
    s.apply(O.Program().add(O.PAULI_ERROR2, 0, 1, p=p).array())
    assert dict(s.terms()) == want


def test_noise_rs_sequence_on_oracle():
    # IZ: IX error (= x(1)) then IY error (= y(1)) returns to +IZ (tests/noise.rs:10-27)
    s = OracleSum(2, 0.0, None)
    s.add_strings(["IZ"], [1.0])
    s.apply(O.Program().add(O.PAULI_ERROR2, 0, 1, p=_p(IX=1.0)).add(O.PAULI_ERROR2, 0, 1, p=_p(IY=1.0)).array())
    assert dict(s.terms()) == {"IZ": 1.0}


def test_depolarize_and_pauli_error_closed_forms_on_oracle():
    # tests/noise.rs:97-175: products of 1 - 4p/3, 1 - 16p/15, and 1 - 2(p_j + p_k)
    s = OracleSum(3, 0.0, None)
    s.add_strings(["ZZZ"], [1.0])
    ps = [0.1, 0.2, 0.3]
    prog = O.Program()
    for q, pq in enumerate(ps):
        prog.add(O.DEPOLARIZE1, q, p=[pq])
    s.apply(prog.array())
    assert abs(s.overlap_with_zero() - np.prod([1 - 4 * p / 3 for p in ps])) < 1e-10
    s = OracleSum(2, 0.0, None)
    s.add_strings(["ZZ"], [1.0])
    s.apply(O.Program().add(O.DEPOLARIZE2, 0, 1, p=[0.1]).array())
    assert abs(s.trace("Z0Z1") - (1 - 16 * 0.1 / 15)) < 1e-10
    px, py, pz = 0.05, 0.10, 0.15
    for w, want in [("I", 1.0), ("X", 1 - 2 * py - 2 * pz), ("Y", 1 - 2 * px - 2 * pz), ("Z", 1 - 2 * px - 2 * py)]:
        s = OracleSum(1, 0.0, None)
        s.add_strings([w], [1.0])
        s.apply(O.Program().add(O.PAULI_ERROR, 0, p=[px, py, pz]).array())
        assert abs(dict(s.terms())[w] - want) < 1e-10


# ---- the same on the HIP engine -------------------------------------------------------------
def _gpu(terms, n, tau=1e-12):
    from ppvm_amd import PauliSum
    return PauliSum(initial_terms=terms, n_qubits=n, min_abs_coeff=tau)


@pytest.mark.gpu
@pytest.mark.parametrize("start,p,want", NOISE_RS_CASES)
def test_noise_rs_on_gpu(start, p, want):
    s = _gpu([start], 2)
    s.two_qubit_pauli_error(0, 1, p=p)
    assert dict(s.terms) == want


@pytest.mark.gpu
def test_noise_rs_sequence_and_rxx_side_on_gpu():
    s = _gpu(["IZ"], 2)
    s.two_qubit_pauli_error(0, 1, p=_p(IX=1.0))
    s.two_qubit_pauli_error(0, 1, p=_p(IY=1.0))
    assert dict(s.terms) == {"IZ": 1.0}
    r = _gpu(["YI"], 2)
    r.rxx(0, 1, theta=math.pi)  # truncates after the call, as the reference's state2
    assert dict(r.terms) == {"YI": -1.0}


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_two_qubit_pauli_error_rows_on_gpu(seed):
    """Every row on the GPU: bit-identical to the oracle's factor and within a
    few ulp of the symplectic closed form, at qubits (1, 3) of a 5-qubit sum."""
    rng = np.random.default_rng(100 + seed)
    p = list(rng.uniform(0.0, 1.0 / 15.0, size=15))
    for w in WORDS:
        full = ["I"] * 5
        full[1], full[3] = w[0], w[1]
        word = "".join(full)
        s = _gpu([word], 5, tau=0.0)
        s.two_qubit_pauli_error(1, 3, p=p)
        (got_w, got), = s.terms
        assert got_w == word
        want = oracle_factor(w, p, 5, 1, 3)
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (w, got, want)
        assert abs(got - closed_form(w, p)) <= 8 * np.finfo(float).eps


@pytest.mark.gpu
def test_depolarize_and_pauli_error_closed_forms_on_gpu():
    s = _gpu(["ZZZ"], 3, tau=0.0)
    ps = [0.1, 0.2, 0.3]
    for q, pq in enumerate(ps):
        s.depolarize1(q, p=pq)
    assert abs(s.overlap_with_zero() - np.prod([1 - 4 * p / 3 for p in ps])) < 1e-10
    s = _gpu(["ZZ"], 2, tau=0.0)
    s.depolarize2(0, 1, p=0.1)
    assert abs(s.trace("Z0Z1") - (1 - 16 * 0.1 / 15)) < 1e-10
    px, py, pz = 0.05, 0.10, 0.15
    for w, want in [("I", 1.0), ("X", 1 - 2 * py - 2 * pz), ("Y", 1 - 2 * px - 2 * pz), ("Z", 1 - 2 * px - 2 * py)]:
        s = _gpu([w], 1, tau=0.0)
        s.pauli_error(0, p=[px, py, pz])
        assert abs(dict(s.terms)[w] - want) < 1e-10


# ---- the reference's golden expectation values on the HIP path -----------------------------
def _trotter_program():
    """crates/ppvm-pauli-sum/tests/trotter.rs:35-54: 4 qubits, 10 steps; per
    step pauli_error then rx(0.1) on every qubit, then per bond pauli_error on
    both qubits and rzz(0.0125), truncating after every call."""
    N, STEPS, TX, TZZ, NOISE = 4, 10, 0.1, 0.0125, [2.5e-5] * 3
    p = O.Program()
    for _ in range(STEPS):
        for i in range(N):
            p.add(O.PAULI_ERROR, i, p=NOISE).truncate()
            p.add(O.RX, i, theta=TX).truncate()
        for i in range(N - 1):
            p.add(O.PAULI_ERROR, i + 1, p=NOISE).truncate()
            p.add(O.PAULI_ERROR, i, p=NOISE).truncate()
            p.add(O.RZZ, i, i + 1, theta=TZZ).truncate()
    return p.array()


@pytest.mark.gpu
@pytest.mark.parametrize("tau", [0.0, 1e-6])
def test_trotter_golden_on_gpu(tau):
    from ppvm_amd._native import Handle
    obs = ["".join("Z" if i == j else "I" for i in range(4)) for j in range(4)]
    h = Handle(4, tau, None)
    x, z = O.pack_strings(obs, 4)
    h.add_words(x, z, np.ones(4))
    prog = _trotter_program()
    h.apply(prog)
    v = h.overlap_with_zero()
    o = OracleSum(4, tau, None)
    o.add_strings(obs)
    o.apply(prog)
    assert abs(v - o.overlap_with_zero()) <= 1e-12
    if tau == 0.0:
        assert abs(v - 2.1610566562692544) < 1e-9  # tests/trotter.rs:104
    else:
        assert abs(v - 2.1610566562692544) < 1e-6  # truncation drift bound, tests/trotter.rs:112-115


@pytest.mark.gpu
def test_cnot_golden_on_gpu():
    """crates/ppvm-pauli-sum/tests/cnot.rs:8-27: rz, ry, rz on both qubits,
    cnot, rx on both, from ZZ; overlap 0.18803675917759355."""
    from ppvm_amd._native import Handle
    h = Handle(2, 0.0, None)
    x, z = O.pack_strings(["ZZ"], 2)
    h.add_words(x, z, np.ones(1))
    p = O.Program()
    for q in (0, 1):
        p.add(O.RZ, q, theta=1.1).add(O.RY, q, theta=2.1).add(O.RZ, q, theta=1.1)
    p.add(O.CNOT, 0, 1).add(O.RX, 0, theta=2.1).add(O.RX, 1, theta=2.1)
    h.apply(p.array())
    v = h.overlap_with_zero()
    assert abs(v - 0.18803675917759355) <= 4 * np.finfo(float).eps * abs(v)


@pytest.mark.gpu
def test_ghz_on_gpu():
    from ppvm_amd._native import Handle
    h = Handle(2, 0.0, None)
    x, z = O.pack_strings(["ZZ"], 2)
    h.add_words(x, z, np.ones(1))
    h.apply(O.Program().add(O.CNOT, 0, 1).add(O.H, 0).array())
    assert h.overlap_with_zero() == 1.0  # tests/ghz.rs:54-70
