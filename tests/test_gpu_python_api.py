"""The Python `PauliSum` drop-in on the GPU, checked the way the reference's
Python tests check theirs (restated, not copied): construction forms and
defaults, string form, copies and equality, weights, overlaps, the gate and
noise table of known conjugations, the per-call `truncate=` cadence, and the
Heisenberg-picture conventions against a small density-matrix simulator.
Restates ppvm-python/test/test_basics.py, test_truncate_kwarg.py and
test_heisenberg_ordering.py.
"""

import functools
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ps(terms, n, coefs=None, **kw):
    from ppvm_amd import PauliSum
    return PauliSum(initial_terms=terms, n_qubits=n, coefficients=coefs or [1.0] * len(terms), **kw)


def _t(s):
    return dict(s.terms)


# ---- construction, display, bookkeeping -------------------------------------------------------
def test_string_form_and_overlap():
    s = _ps(["ZZ"], 2)
    s.cnot(0, 1)
    s.h(0)
    assert str(s) == "1.000 * IZ"
    assert s.overlap_with_zero() == 1.0
    # terms joined by " + ", stably sorted by weight (sum/display.rs:29-44)
    assert str(_ps(["ZX", "IY", "XI"], 2, [0.5, -0.25, 2.0])) == "-0.250 * IY + 2.000 * XI + 0.500 * ZX"


def test_wide_state_with_weight_cap():
    n = 200
    s = _ps(["".join("Z" if i == j else "I" for i in range(n)) for j in range(n)], n, max_pauli_weight=80)
    for i in reversed(range(1, n)):
        s.cnot(i - 1, i)
    s.h(0)
    assert s.overlap_with_zero() == 0.0


def test_copy_equality_and_terms():
    s = _ps(["ZZ"], 2)
    s.cnot(0, 1)
    s.h(0)
    c = s.copy()
    assert c == s and len(c) == len(s) == 1
    assert c.terms == [("IZ", 1.0)]
    c.x(1)
    assert c != s


def test_weights_and_max_weight():
    s = _ps(["IZ"], 2)
    assert s.current_max_weight() == 1
    s.cnot(0, 1)
    s.h(0)
    assert s.current_max_weight() == 2
    w = sorted(_ps(["ZX", "IY"], 2).weights(), key=lambda x: x[1])
    assert w == [("IY", 1), ("ZX", 2)]


def test_overlap_between_sums():
    s = _ps(["IZ"], 2)
    assert s.overlap(s) == 1.0
    assert s.overlap(_ps(["IX"], 2)) == 0.0


def test_new_forms():
    from ppvm_amd import PauliSum
    assert PauliSum.new(2, "IX").terms == [("IX", 1.0)]
    assert PauliSum.new(2, ("IX", 0.25)).terms == [("IX", 0.25)]
    assert PauliSum.new(2, "X1") == PauliSum.new(2, "IX")
    s = PauliSum.new(3, [("Y1", 0.1), "ZIZ"])
    assert set(s.terms) == {("IYI", 0.1), ("ZIZ", 1.0)}
    n = 17
    s = PauliSum.new(n, [f"Z{i}" for i in range(n)])
    assert set(s.terms) == {("".join("Z" if i == j else "I" for i in range(n)), 1.0) for j in range(n)}
    with pytest.raises(ValueError, match="out of range"):
        PauliSum.new(2, "X2")


# ---- gate and noise table -------------------------------------------------------------------------
@pytest.mark.parametrize("start,gate,args,want", [
    ("XI", "x", (0,), {"XI": 1.0}),
    ("ZI", "x", (0,), {"ZI": -1.0}),
    ("IY", "y", (1,), {"IY": 1.0}),
    ("IZ", "z", (1,), {"IZ": 1.0}),
    ("ZI", "h", (0,), {"XI": 1.0}),
    ("XI", "h", (0,), {"ZI": 1.0}),
    ("XI", "s", (0,), {"YI": -1.0}),
    ("YI", "s_dag", (0,), {"XI": -1.0}),
    ("IZ", "cnot", (0, 1), {"ZZ": 1.0}),
    ("IX", "cnot", (0, 1), {"IX": 1.0}),
    ("IX", "cz", (0, 1), {"ZX": 1.0}),
    ("XI", "cz", (0, 1), {"XZ": 1.0}),
])
def test_clifford_table(start, gate, args, want):
    s = _ps([start], 2)
    getattr(s, gate)(*args)
    assert _t(s) == want


@pytest.mark.parametrize("start,call,word,value", [
    ("ZI", lambda s: s.ry(0, theta=-math.pi / 2), "XI", 1.0),
    ("ZI", lambda s: s.rx(0, theta=math.pi / 2), "YI", 1.0),
    ("XI", lambda s: s.rz(0, theta=-math.pi / 2), "YI", 1.0),
    ("ZI", lambda s: s.r(0, 0.0, math.pi / 2), "YI", 1.0),
    ("ZI", lambda s: s.r(0, math.pi / 2, math.pi / 2), "XI", -1.0),
])
def test_rotation_table(start, call, word, value):
    s = _ps([start], 2)
    call(s)
    assert _t(s).get(word, 0.0) == pytest.approx(value)


def test_noise_table():
    s = _ps(["ZI"], 2)
    s.pauli_error(0, p=[0.0, 0.0, 0.5])
    assert _t(s) == {"ZI": 1.0}
    s = _ps(["XI"], 2)
    s.pauli_error(0, p=[0.0, 0.0, 0.5])
    assert len(s) == 0
    s = _ps(["ZI"], 2)
    s.pauli_error(0, p=[0.25, 0.25, 0.25])
    assert len(s) == 0
    p = [0.0] * 15
    p[0] = 1.0  # IX anticommutes with IZ
    s = _ps(["IZ"], 2)
    s.two_qubit_pauli_error(0, 1, p=p)
    assert _t(s)["IZ"] == pytest.approx(-1.0)
    s = _ps(["ZI"], 2)
    s.depolarize1(0, p=0.1)
    assert _t(s)["ZI"] == pytest.approx(1.0 - 4 * 0.1 / 3)
    s = _ps(["ZI"], 2)
    s.depolarize1(0, p=0.75)
    assert len(s) == 0
    s = _ps(["ZZ"], 2)
    s.depolarize2(0, 1, p=0.1)
    assert _t(s)["ZZ"] == pytest.approx(1.0 - 16 * 0.1 / 15)
    s = _ps(["XI"], 2)
    s.amplitude_damping(0, 0.2)
    assert _t(s) == {"XI": pytest.approx(0.8 ** 0.5)}
    s = _ps(["ZI"], 2)
    s.amplitude_damping(1, 0.2)
    assert _t(s) == {"ZI": 1.0}


# ---- truncate= cadence ----------------------------------------------------------------------------
def test_default_truncate_equals_explicit_true():
    a = _ps(["IZI"], 3, min_abs_coeff=0.05)
    b = _ps(["IZI"], 3, min_abs_coeff=0.05)
    a.rxx(0, 1, theta=0.5)
    a.ryy(0, 1, theta=0.5)
    b.rxx(0, 1, theta=0.5, truncate=True)
    b.ryy(0, 1, theta=0.5, truncate=True)
    assert _t(a) == _t(b)


def test_deferred_truncate_keeps_intermediate_terms():
    plain = _ps(["IZI"], 3, min_abs_coeff=0.5)
    plain.rxx(0, 1, theta=0.4)
    assert _t(plain) == {"IZI": math.cos(0.4)}
    deferred = _ps(["IZI"], 3, min_abs_coeff=0.5)
    deferred.rxx(0, 1, theta=0.4, truncate=False)
    assert "IZI" in _t(deferred) and len(deferred) > 1
    deferred.truncate()
    assert _t(deferred) == _t(plain)


def test_deferred_pair_then_truncate_equals_truncate_on_last():
    a = _ps(["IZIZ"], 4, min_abs_coeff=1e-2)
    b = _ps(["IZIZ"], 4, min_abs_coeff=1e-2)
    a.rxx(1, 2, theta=0.4, truncate=False)
    a.ryy(1, 2, theta=0.4, truncate=False)
    a.truncate()
    b.rxx(1, 2, theta=0.4, truncate=False)
    b.ryy(1, 2, theta=0.4)
    assert _t(a) == _t(b)


def test_explicit_truncate_and_noise_kwarg():
    s = _ps(["ZI", "XI"], 2, [0.5, 1e-4], min_abs_coeff=1e-3)
    assert "XI" in _t(s)  # the constructor does not truncate
    s.truncate()
    assert _t(s) == {"ZI": 0.5}
    a = _ps(["ZZZ"], 3, min_abs_coeff=1e-2)
    b = _ps(["ZZZ"], 3, min_abs_coeff=1e-2)
    a.pauli_error(1, p=[0.0, 0.0, 0.3], truncate=False)
    a.pauli_error(2, p=[0.0, 0.0, 0.3], truncate=False)
    a.truncate()
    b.pauli_error(1, p=[0.0, 0.0, 0.3])
    b.pauli_error(2, p=[0.0, 0.0, 0.3])
    assert _t(a) == _t(b)


# ---- Heisenberg-picture conventions vs a density-matrix simulation -------------------------------
_I2 = np.eye(2, dtype=complex)
_P = {"I": _I2, "X": np.array([[0, 1], [1, 0]], dtype=complex),
      "Y": np.array([[0, -1j], [1j, 0]], dtype=complex), "Z": np.diag([1.0 + 0j, -1.0])}
_H = np.array([[1, 1], [1, -1]], dtype=complex) / math.sqrt(2)
_S = np.diag([1.0 + 0j, 1j])


def _on(op, q, n):
    return functools.reduce(np.kron, [op if i == q else _I2 for i in range(n)])


def _cnot(c, t, n):
    p0 = np.diag([1.0 + 0j, 0.0])
    p1 = np.diag([0.0 + 0j, 1.0])
    return _on(p0, c, n) + _on(p1, c, n) @ _on(_P["X"], t, n)


def _rot(P, theta):
    return math.cos(theta / 2) * np.eye(P.shape[0]) - 1j * math.sin(theta / 2) * P


def _forward(ops, n):
    """rho after the circuit (gates as unitaries, noise as Kraus maps), from |0..0>."""
    rho = np.zeros((2 ** n, 2 ** n), dtype=complex)
    rho[0, 0] = 1.0
    for op in ops:
        kind = op[0]
        if kind in ("h", "s", "cnot", "rx", "rz", "rzz"):
            if kind == "h":
                U = _on(_H, op[1], n)
            elif kind == "s":
                U = _on(_S, op[1], n)
            elif kind == "cnot":
                U = _cnot(op[1], op[2], n)
            elif kind == "rx":
                U = _rot(_on(_P["X"], op[1], n), op[2])
            elif kind == "rz":
                U = _rot(_on(_P["Z"], op[1], n), op[2])
            else:
                U = _rot(_on(_P["Z"], op[1], n) @ _on(_P["Z"], op[2], n), op[3])
            rho = U @ rho @ U.conj().T
        elif kind == "depolarize1":
            q, p = op[1], op[2]
            rho = (1 - p) * rho + sum((p / 3) * _on(_P[a], q, n) @ rho @ _on(_P[a], q, n) for a in "XYZ")
        elif kind == "amplitude_damping":
            q, g = op[1], op[2]
            K0 = _on(np.array([[1, 0], [0, math.sqrt(1 - g)]], dtype=complex), q, n)
            K1 = _on(np.array([[0, math.sqrt(g)], [0, 0]], dtype=complex), q, n)
            rho = K0 @ rho @ K0.conj().T + K1 @ rho @ K1.conj().T
    return rho


def _backward(ops, n, obs):
    s = _ps([obs], n, min_abs_coeff=0.0)
    for op in reversed(ops):  # reverse the order, keep the angles
        kind = op[0]
        if kind in ("h", "s"):
            getattr(s, kind)(op[1])
        elif kind == "cnot":
            s.cnot(op[1], op[2])
        elif kind in ("rx", "rz"):
            getattr(s, kind)(op[1], theta=op[2])
        elif kind == "rzz":
            s.rzz(op[1], op[2], theta=op[3])
        elif kind == "depolarize1":
            s.depolarize1(op[1], p=op[2])
        else:
            s.amplitude_damping(op[1], op[2])
    return s.overlap_with_zero()


def _circuit(seed, n):
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(14):
        k = rng.integers(0, 8)
        a, b = (int(v) for v in rng.choice(n, 2, replace=False))
        th = float(rng.uniform(-2, 2))
        ops.append([("h", a), ("s", a), ("cnot", a, b), ("rx", a, th), ("rz", a, th), ("rzz", a, b, th),
                    ("depolarize1", a, 0.05), ("amplitude_damping", a, 0.3)][k])
    return ops


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_backward_pauli_sum_matches_density_matrix(seed):
    n = 3
    ops = _circuit(seed, n)
    rho = _forward(ops, n)
    for obs in ("ZII", "IZZ", "XIY"):
        O = functools.reduce(np.kron, [_P[ch] for ch in obs])
        want = float(np.trace(O @ rho).real)
        assert abs(_backward(ops, n, obs) - want) < 1e-10


def test_negated_angles_give_the_inverse_circuit():
    """Running backwards reverses the order only; negating the angles as well
    simulates a different circuit (here <Y_0> flips sign)."""
    n = 3
    ops = [("rx", 0, 0.7), ("rzz", 0, 1, 0.9), ("rx", 1, 0.5), ("rz", 0, -1.1), ("rx", 0, 0.4), ("cnot", 1, 2)]
    O = functools.reduce(np.kron, [_P[ch] for ch in "YII"])
    want = float(np.trace(O @ _forward(ops, n)).real)
    neg = [(o[0], *o[1:-1], -o[-1]) if o[0] in ("rx", "rz", "rzz") else o for o in ops]
    assert abs(_backward(ops, n, "YII") - want) < 1e-10
    assert abs(_backward(neg, n, "YII") - want) > 1.0
