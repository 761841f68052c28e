"""ppvm_amd.core: the ``ppvm._core`` class family on the GPU engine.

CPU: the 32 class names, the method surface of ppvm-python/src/ppvm/_core.pyi:3-58,
and the tier the reference wrapper picks (paulisum.py:176-182).  GPU: every
gate and channel called with the native module's signature (targets list,
angle / probabilities, truncate=True) gives the terms, order and coefficient
bits of the oracle running the same ops; the truncate kwarg, the stim
aliases, odd pair lists, tier capacity, equality, copies and display.
"""

import copy
import inspect
import math

import numpy as np
import pytest

from oracle.pyoracle import LossyOracleSum, OracleSum
from ppvm_amd import core
from ppvm_amd import ops as O

PYI_METHODS = ["trace", "overlap_with_zero", "overlap", "truncate", "x", "y", "z", "h", "s", "cnot", "cx", "zcx",
               "cz", "zcz", "cy", "zcy", "s_dag", "sqrt_x", "sqrt_y", "sqrt_x_dag", "sqrt_y_dag", "rx", "ry", "rz",
               "rxx", "ryy", "rzz", "x_error", "y_error", "z_error", "pauli_error", "two_qubit_pauli_error",
               "depolarize1", "depolarize2", "amplitude_damping", "terms", "weights", "current_max_weight",
               "__copy__", "__deepcopy__", "__len__", "__str__", "__repr__"]
LOSS_METHODS = ["loss_channel", "correlated_loss_channel", "reset_loss_channel"]


def test_class_family_and_surface():
    for k in range(16):
        plain = getattr(core, f"PauliSumIndexMapFxHash{k}")
        lossy = getattr(core, f"PauliSumLossIndexMapFxHash{k}")
        assert plain._tier_bytes == lossy._tier_bytes == 2 ** k
        assert issubclass(lossy, plain.__mro__[1])
        for m in PYI_METHODS:
            assert callable(getattr(plain, m)), m
        for m in LOSS_METHODS:
            assert callable(getattr(lossy, m)) and not hasattr(plain, m), m
    sig = inspect.signature(core.PauliSumIndexMapFxHash0.__init__)
    assert list(sig.parameters)[1:] == ["n_qubits", "min_abs_coeff", "max_pauli_weight", "max_loss_weight",
                                        "terms", "coefficients", "preserve_strings"]
    assert sig.parameters["min_abs_coeff"].default == 1e-10
    for m in ("rx", "rzz", "pauli_error", "depolarize2"):
        p = inspect.signature(getattr(core.PauliSumIndexMapFxHash0, m)).parameters
        assert list(p)[1] == "targets" and p["truncate"].default is True


def test_tier_selection_matches_the_reference_formula():
    # paulisum.py:176-182: N = ceil(n / 8) bytes, the first k with 2**k > N
    for n in range(1, 2048):
        N = math.ceil(n / 8.0)
        want = next(k for k in range(16) if 2 ** k > N)
        assert core.tier_for(n) == want
        assert n <= 8 * 2 ** want  # the chosen tier can hold the qubits
    assert core.tier_for(8) == 1 and core.tier_for(9) == 2 and core.tier_for(100) == 4


# ---- GPU ---------------------------------------------------------------------------------
def _cls(n, lossy=False):
    name = "PauliSumLossIndexMapFxHash" if lossy else "PauliSumIndexMapFxHash"
    return getattr(core, f"{name}{core.tier_for(n)}")


CALLS = [  # (method, args, op rows for the oracle)
    ("h", ([0, 2],), [(O.H, 0), (O.H, 2)]),
    ("s", ([1],), [(O.S, 1)]),
    ("s_dag", ([3],), [(O.S_DAG, 3)]),
    ("sqrt_x", ([2],), [(O.SQRT_X, 2)]),
    ("sqrt_y_dag", ([0],), [(O.SQRT_Y_DAG, 0)]),
    ("cnot", ([0, 1, 2, 3],), [(O.CNOT, 0, 1), (O.CNOT, 2, 3)]),
    ("zcz", ([1, 2],), [(O.CZ, 1, 2)]),
    ("cx", ([3, 0],), [(O.CNOT, 3, 0)]),
    ("cy", ([2, 0],), [(O.CY, 2, 0)]),
    ("rx", ([0, 1, 2, 3], 0.37), [(O.RX, q, 0, 0.37) for q in range(4)]),
    ("rzz", ([0, 1, 1, 2], -0.81), [(O.RZZ, 0, 1, -0.81), (O.RZZ, 1, 2, -0.81)]),
    ("ryy", ([2, 3], 1.3), [(O.RYY, 2, 3, 1.3)]),
    ("ry", ([1], 0.5), [(O.RY, 1, 0, 0.5)]),
    ("rxx", ([3, 1], 0.21), [(O.RXX, 3, 1, 0.21)]),
    ("rz", ([0, 3], 2.2), [(O.RZ, 0, 0, 2.2), (O.RZ, 3, 0, 2.2)]),
    ("x_error", ([0], 0.03), [(O.PAULI_ERROR, 0, 0, 0.0, [0.03, 0, 0])]),
    ("pauli_error", ([1, 2], [0.01, 0.02, 0.03]), [(O.PAULI_ERROR, q, 0, 0.0, [0.01, 0.02, 0.03]) for q in (1, 2)]),
    ("depolarize1", ([3], 0.05), [(O.DEPOLARIZE1, 3, 0, 0.0, [0.05])]),
    ("depolarize2", ([0, 1], 0.04), [(O.DEPOLARIZE2, 0, 1, 0.0, [0.04])]),
    ("two_qubit_pauli_error", ([2, 3], [0.004 * (i + 1) for i in range(15)]),
     [(O.PAULI_ERROR2, 2, 3, 0.0, [0.004 * (i + 1) for i in range(15)])]),
    ("amplitude_damping", (1, 0.2), [(O.AMPLITUDE_DAMPING, 1, 0, 0.0, [0.2])]),
]


def _oracle_apply(o, rows, truncate=True):
    p = O.Program()
    for r in rows:
        p.add(*r[:4], p=r[4] if len(r) > 4 else None)
    if truncate:
        p.truncate()
    o.apply(p.array())


def _same(g, o, ctx):
    gt, ot = g.terms(), o.terms()
    assert [k for k, _ in gt] == [k for k, _ in ot], f"{ctx}: words or order differ"
    gb = np.array([v for _, v in gt]).view(np.uint64)
    ob = np.array([v for _, v in ot]).view(np.uint64)
    assert np.array_equal(gb, ob), f"{ctx}: coefficient bits differ"


@pytest.mark.gpu
@pytest.mark.parametrize("tau,maxw", [(1e-10, None), (1e-3, 3)])
def test_every_method_matches_the_oracle(tau, maxw):
    n = 4
    terms, coefs = ["ZIII", "XYIZ", "IZZI", "YIXX"], [1.0, -0.5, 0.25, 0.75]
    kw = {} if maxw is None else {"max_pauli_weight": maxw}
    g = _cls(n)(n, tau, terms=terms, coefficients=coefs, **kw)
    o = OracleSum(n, tau, maxw)
    o.add_strings(terms, coefs)
    for name, args, rows in CALLS:
        getattr(g, name)(*args)
        _oracle_apply(o, rows)
        _same(g, o, name)
    assert abs(g.overlap_with_zero() - o.overlap_with_zero()) <= 1e-12
    assert abs(g.trace("Z?*") - o.trace("Z?*")) <= 1e-12
    assert abs(g.trace("[XY]0Z?*") - o.trace("[XY]0Z?*")) <= 1e-12
    assert len(g) == len(o)


@pytest.mark.gpu
def test_truncate_kwarg_defers_the_cut():
    """rx(0.04) twice under a 0.05 cut: truncating after each call drops the
    first child (sin 0.04 < 0.05); deferring keeps it until it merges."""
    n = 3
    a = _cls(n)(n, 0.05, terms=["ZII"], coefficients=[1.0])
    b = _cls(n)(n, 0.05, terms=["ZII"], coefficients=[1.0])
    a.rx([0], 0.04)
    a.rx([0], 0.04)
    b.rx([0], 0.04, truncate=False)
    b.rx([0], 0.04, truncate=False)
    b.truncate()
    oa = OracleSum(n, 0.05, None)
    oa.add_strings(["ZII"])
    _oracle_apply(oa, [(O.RX, 0, 0, 0.04)])
    _oracle_apply(oa, [(O.RX, 0, 0, 0.04)])
    ob = OracleSum(n, 0.05, None)
    ob.add_strings(["ZII"])
    _oracle_apply(ob, [(O.RX, 0, 0, 0.04), (O.RX, 0, 0, 0.04)])
    _same(a, oa, "truncate=True")
    _same(b, ob, "truncate=False")
    assert a != b


@pytest.mark.gpu
def test_errors_equality_copies_and_display():
    n = 5
    g = _cls(n)(n, terms=["ZZIII", "IXIII"], coefficients=[0.5, -0.25])
    with pytest.raises(ValueError, match="even number of targets"):
        g.cnot([0, 1, 2])
    with pytest.raises(ValueError, match="even number of targets"):
        g.rzz([0], 0.1)
    with pytest.raises(ValueError, match="stores 8 qubits"):
        core.PauliSumIndexMapFxHash0(9)
    assert core.PauliSumIndexMapFxHash0(8, terms=["IIIIIIIZ"], coefficients=[1.0]).overlap_with_zero() == 1.0
    h = copy.copy(g)
    d = copy.deepcopy(g)
    assert h == g and d == g and not (h != g)
    h.h([0])
    assert h != g and d == g
    with pytest.raises(NotImplementedError):
        _ = g < h
    assert str(g) == "-0.250 * IXIII + 0.500 * ZZIII"
    assert repr(g) == "PauliSum(-0.250 * IXIII + 0.500 * ZZIII)"
    assert g.weights() == [("ZZIII", 2), ("IXIII", 1)] and g.current_max_weight() == 2
    assert abs(g.overlap(h) - 0.0625) < 1e-15 and abs(g.overlap(d) - (0.25 + 0.0625)) < 1e-15
    e = _cls(n)(n)  # no terms is allowed
    assert len(e) == 0 and e.current_max_weight() == 0 and str(e) == ""


@pytest.mark.gpu
def test_lossy_class_loss_channels_match_the_oracle():
    n = 4
    g = _cls(n, lossy=True)(n, 1e-10, max_loss_weight=2, terms=["ZLIZ", "LLXI", "IZZI"],
                            coefficients=[1.0, 0.5, -0.75])
    o = LossyOracleSum(n, 1e-10, None, 2)
    o.add_strings(["ZLIZ", "LLXI", "IZZI"], [1.0, 0.5, -0.75])
    g.reset_loss_channel(2)
    _oracle_apply(o, [(O.RESET_LOSS, 2)])
    g.loss_channel(0, 0.1)
    _oracle_apply(o, [(O.LOSS, 0, 0, 0.0, [0.1])])
    g.correlated_loss_channel(0, 1, [0.2, 0.1, 0.3])
    _oracle_apply(o, [(O.CORRELATED_LOSS, 0, 1, 0.0, [0.2, 0.1, 0.3])])
    g.cnot([3, 2])
    _oracle_apply(o, [(O.CNOT, 3, 2)])
    _same(g, o, "lossy")
    assert abs(g.overlap_with_zero() - o.overlap_with_zero()) <= 1e-12
