"""GPU engine vs CPU oracle: bit-identical term sets and coefficients.

Every test drives the HIP library through the C ABI (via the ctypes shim) and
replays the same op program on the oracle.  Bar: identical keys, identical
coefficient bits and identical (insertion) order; overlap_with_zero within
1e-10 absolute (the GPU sums in a different association).
"""

import numpy as np
import pytest

from ppvm_amd import ops as O
from ppvm_amd import workloads as WL
from tests.helpers import (assert_same_terms, gpu_for, oracle_for, random_program, random_terms,
                           run_segments)

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["global", "partition", "partition_sort", "small"])
def pairing_path(request, monkeypatch):
    """Run every test on each execution path: the arena pipeline with the
    single global-table pairing kernel, the arena pipeline with the radix
    partition + LDS pairing kernels (forced at every size), the same with the
    sort-based pairing inside the buckets, and the default settings, under
    which a small sum runs its steps in LDS in one workgroup (k_small) and
    hands over to the arena pipeline when it outgrows LDS.  Handles read the
    choice when they are created."""
    monkeypatch.delenv("PPVM_PAIR_METHOD", raising=False)
    if request.param == "small":
        monkeypatch.delenv("PPVM_PARTITION_MIN_ENTRIES", raising=False)
        monkeypatch.delenv("PPVM_SMALL_MAX_ENTRIES", raising=False)
    else:
        monkeypatch.setenv("PPVM_SMALL_MAX_ENTRIES", "0")
        monkeypatch.setenv("PPVM_PARTITION_MIN_ENTRIES",
                           str(2**63) if request.param == "global" else "0")
        if request.param == "partition_sort":
            monkeypatch.setenv("PPVM_PAIR_METHOD", "sort")
    return request.param


def _pair(n, strings, coeffs, tau=0.0, maxw=None, cap=0):
    from oracle.pyoracle import OracleSum
    from ppvm_amd import PauliSum
    g = PauliSum(strings, n, coefficients=coeffs, min_abs_coeff=tau, max_pauli_weight=maxw,
                 capacity_hint=cap)
    r = OracleSum(n, tau, maxw)
    r.add_strings(strings, coeffs)
    return g, r


def test_device_present():
    from ppvm_amd import _native
    info = _native.device_info()
    print(info)
    assert "gfx950" in info


@pytest.mark.parametrize("n", [5, 70, 130, 300])
def test_every_gate_on_random_terms(n):
    """Each gate kind on 200 random terms, no truncation (W = 1, 2, 4, 8)."""
    rng = np.random.default_rng(1000 + n)
    strings = random_terms(rng, n, 200)
    coeffs = rng.normal(size=200)
    for kind in ["x", "y", "z", "h", "s", "s_dag", "sqrt_x", "sqrt_x_dag", "sqrt_y", "sqrt_y_dag",
                 "cnot", "cz", "cy", "rx", "ry", "rz", "rxx", "ryy", "rzz"]:
        g, r = _pair(n, strings, coeffs)
        prog = random_program(rng, n, 6, trunc_every=False, kinds=[kind])
        g.apply(prog)
        r.apply(prog)
        assert_same_terms(g._export(), r.export(), True, f"n={n} gate={kind}")


@pytest.mark.parametrize("n,tau,maxw", [(6, 1e-3, None), (40, 1e-2, 6), (100, 1e-4, 10)])
def test_random_circuits_truncated(n, tau, maxw):
    rng = np.random.default_rng(7 * n)
    strings = random_terms(rng, n, 20)
    coeffs = rng.normal(size=20)
    g, r = _pair(n, strings, coeffs, tau, maxw)
    prog = random_program(rng, n, 120, trunc_every=True)
    marks = list(range(1, len(prog), 20))
    run_segments(g, r, prog, marks, ctx=f"random n={n}")


def test_chain_tfim_every_layer():
    w = WL.chain_tfim()
    g, r = gpu_for(w), oracle_for(w)
    run_segments(g, r, w.program, w.layer_marks, ctx="chain_tfim")
    assert g.stats()["term_updates"] == r.term_updates()


def test_lattice_prefix_every_layer():
    w = WL.truncate_steps(WL.lattice_tfim(), 5)
    g, r = gpu_for(w), oracle_for(w)
    run_segments(g, r, w.program, w.layer_marks, ctx="lattice_tfim")
    assert g.stats()["term_updates"] == r.term_updates()


def test_random_circuit_every_layer():
    w = WL.random_circuit()
    g, r = gpu_for(w), oracle_for(w)
    run_segments(g, r, w.program, w.layer_marks, ctx="random_circuit")


def test_growth_from_tiny_capacity():
    """Arena overflow is detected on the device and replayed after growth."""
    w = WL.truncate_steps(WL.lattice_tfim(), 4)
    g, r = gpu_for(w, capacity_hint=1), oracle_for(w)
    g.apply(w.program)
    r.apply(w.program)
    assert_same_terms(g._export(), r.export(), True, "growth")
    st = g.stats()
    assert st["regrows"] >= 1


def test_deferred_truncation_cadence():
    """Per-call cadence: truncate only after each layer (Python *_many calls)."""
    w = WL.chain_tfim(n=16, steps=4)
    prog = w.program[w.program["code"] != O.TRUNCATE]
    p2 = O.Program()
    for i, op in enumerate(prog):
        p2.add(int(op["code"]), int(op["q0"]), int(op["q1"]), float(op["theta"]))
        if i % 16 == 15:
            p2.truncate()
    arr = p2.array()
    g, r = gpu_for(w), oracle_for(w)
    g.apply(arr)
    r.apply(arr)
    assert_same_terms(g._export(), r.export(), True, "deferred")


def test_noise_channels():
    rng = np.random.default_rng(5)
    n = 9
    strings = random_terms(rng, n, 60)
    coeffs = rng.normal(size=60)
    g, r = _pair(n, strings, coeffs, 1e-3)
    p = O.Program()
    for _ in range(20):
        a, b = rng.choice(n, size=2, replace=False)
        p.add(O.PAULI_ERROR, int(a), p=rng.uniform(0, 0.05, 3)).truncate()
        p.add(O.DEPOLARIZE1, int(b), p=[0.03]).truncate()
        p.add(O.DEPOLARIZE2, int(a), int(b), p=[0.02]).truncate()
        p.add(O.PAULI_ERROR2, int(b), int(a), p=rng.uniform(0, 0.01, 15)).truncate()
        p.add(O.RX, int(a), theta=0.3).truncate()
    g.apply(p.array())
    r.apply(p.array())
    assert_same_terms(g._export(), r.export(), True, "noise")


def test_preserve_strings_kept():
    n = 4
    g, r = _pair(n, ["ZIII", "XIII"], [1.0, 1e-9], tau=1e-6)
    g2 = type(g)(["ZIII", "XIII"], n, coefficients=[1.0, 1e-9], min_abs_coeff=1e-6,
                 preserve_strings=["XIII"])
    g2.truncate()
    assert dict(g2.terms) == {"ZIII": 1.0, "XIII": 1e-9}
    g.truncate()
    assert dict(g.terms) == {"ZIII": 1.0}


def _low_weight_terms(rng, n, act, count):
    """count distinct strings of weight 1-3 on the last act of n qubits."""
    out = set()
    while len(out) < count:
        w = ["I"] * n
        for q in rng.choice(act, size=int(rng.integers(1, 4)), replace=False):
            w[n - act + int(q)] = "XYZ"[int(rng.integers(0, 3))]
        out.add("".join(w))
    return sorted(out)


def _dropped_keys(n, strings, coeffs, tau, maxw, prog, count):
    """Keys that a run of ``prog`` without a preserve set truncates away, in the
    order they first vanish (the oracle, op by op)."""
    from oracle.pyoracle import OracleSum
    r = OracleSum(n, tau, maxw)
    r.add_strings(strings, coeffs)
    prev, out = set(strings), []
    for i in range(len(prog)):
        r.apply(prog[i:i + 1])
        cur = {k for k, _ in r.terms()}
        out += [k for k in sorted(prev - cur) if k not in out]
        prev = cur
        if len(out) >= count:
            break
    return out[:count]


@pytest.mark.parametrize("n,tau,maxw,every,cap", [(6, 2e-2, None, 1, 0), (40, 1e-2, 5, 1, 0),
                                                  (70, 5e-3, 6, 5, 0), (40, 1e-2, 5, 1, 1)])
def test_preserved_strings_follow_the_reference_order(n, tau, maxw, every, cap):
    """A preserved string that a truncation would drop leaves its place and
    follows every other term, after every truncation (sum/data.rs:283-310):
    through rotations, Cliffords under a weight cap, truncation every op or
    every fifth op, and arena growth with replay (cap 1)."""
    from oracle.pyoracle import OracleSum
    from ppvm_amd import PauliSum
    rng = np.random.default_rng(40 + n + cap)
    act = min(n, 8)  # the program acts on the last 8 qubits (word 0 and, for n > 64, word 1)
    strings = _low_weight_terms(rng, n, act, 12)
    coeffs = rng.normal(size=12)
    raw = random_program(rng, act, 150, trunc_every=False)
    raw["q0"] += n - act
    two = np.isin(raw["code"], list(O.TWO_QUBIT))
    raw["q1"][two] += n - act
    p = O.Program()
    for i, op in enumerate(raw):
        p.add(int(op["code"]), int(op["q0"]), int(op["q1"]), float(op["theta"]))
        if i % every == every - 1:
            p.truncate()
    prog = p.array()
    keep = _dropped_keys(n, strings, coeffs, tau, maxw, prog, 8)
    assert len(keep) >= 4
    g = PauliSum(strings, n, coefficients=coeffs, min_abs_coeff=tau, max_pauli_weight=maxw,
                 preserve_strings=keep, capacity_hint=cap)
    r = OracleSum(n, tau, maxw)
    r.add_strings(strings, coeffs)
    r.set_preserve(keep)
    rescued, prev = 0, 0
    for m in list(range(2 * every - 1, len(prog), 12)) + [len(prog) - 1]:
        g.apply(prog[prev:m + 1])
        r.apply(prog[prev:m + 1])
        prev = m + 1
        assert_same_terms(g._export(), r.export(), True, f"preserve n={n} op {m}")
        # preserved entries present only because they are preserved
        rescued += sum(1 for k, c in r.terms() if k in keep and
                       (abs(c) < tau or (maxw is not None and sum(ch != "I" for ch in k) > maxw)))
    assert rescued >= 2  # the set is exercised


def test_clone_is_deep():
    w = WL.truncate_steps(WL.chain_tfim(), 2)
    g = gpu_for(w)
    g.apply(w.program)
    c = g.copy()
    g.rx(3, 0.7)
    r = oracle_for(w)
    r.apply(w.program)
    assert_same_terms(c._export(), r.export(), True, "clone")


@pytest.mark.parametrize("n,seed,nq_local", [(12, 1, 4), (70, 2, 8), (130, 3, 10)])
def test_sparse_support_every_gate_kind(n, seed, nq_local):
    """One weight-1 observable and every gate kind on a few qubits near it, then
    on all qubits: most rotations act outside the terms' support, which the
    engine drops without launching the rotation kernels.  Cliffords move the
    support (H swaps x and z, CNOT/CZ/CY spread it), so a support estimate that
    is too narrow would drop a rotation that acts and break parity here."""
    rng = np.random.default_rng(seed)
    q0 = int(rng.integers(0, n - nq_local))
    word = ["I"] * n
    word[q0] = "Z"
    g, r = _pair(n, ["".join(word)], [1.0], tau=1e-7)
    local = random_program(rng, nq_local, 150, trunc_every=True)
    for f in ("q0", "q1"):
        local[f] += q0
    wide = random_program(rng, n, 150, trunc_every=True)
    prog = np.concatenate([local, wide])
    run_segments(g, r, prog, list(range(9, len(prog), 20)), ctx=f"sparse support n={n}")
    assert g.stats()["term_updates"] == r.term_updates()


def test_rotation_outside_support_is_dropped():
    """A rotation whose generator meets no term's support launches no rotation
    kernel (no entry scanned) but still counts its term updates."""
    from ppvm_amd import PauliSum
    g = PauliSum(["ZIIII"], 5, min_abs_coeff=1e-9)
    g.rz(3, theta=0.4)     # Z generator, no X/Y anywhere: commutes with every term
    g.rzz(2, 4, theta=0.3)
    c = g._h.counters()
    assert c["rotations"] == 2 and c["scanned"] == 0 and c["term_updates"] == 2
    g.rx(0, theta=0.5)     # anticommutes with Z_0: runs
    c = g._h.counters()
    assert c["rotations"] == 3 and c["scanned"] == 1 and len(g) == 2
    g.rz(0, theta=0.2)     # the child Y_0 now has x set on qubit 0: runs
    c = g._h.counters()
    assert c["scanned"] == 3 and c["wl_total"] == 2


def test_map_ops_on_identity_qubits_are_dropped():
    """A Clifford or noise op on qubits where every term holds the identity
    changes nothing: it is dropped (no map pass) but still counts as a gate
    and books its term updates, also between the truncation points of a map
    program that does run."""
    from ppvm_amd import PauliSum
    g = PauliSum(["ZIIII", "XIIII"], 5, min_abs_coeff=1e-9)
    g.truncate()  # the sum as a truncation leaves it
    b = g._h.counters()
    g.h(3)
    g.cnot(2, 4)
    g.pauli_error(3, p=[0.01, 0.02, 0.03])
    c = g._h.counters()
    assert c["gates"] - b["gates"] == 3 and c["term_updates"] - b["term_updates"] == 6
    assert c["map_launches"] == b["map_launches"]
    g.h(0)          # acts on both terms
    g.s(2)          # dropped, after a truncation point of the same program
    g.cnot(0, 1)    # acts
    c = g._h.counters()
    assert c["gates"] - b["gates"] == 6 and c["term_updates"] - b["term_updates"] == 12
    assert dict(g.terms) == {"XXIII": 1.0, "ZIIII": 1.0}


def test_determinism():
    w = WL.truncate_steps(WL.lattice_tfim(), 4)
    a, b = gpu_for(w), gpu_for(w)
    a.apply(w.program)
    b.apply(w.program)
    ea, eb = a._export(), b._export()
    assert all(np.array_equal(u, v) for u, v in zip(ea, eb))
    assert a.overlap_with_zero() == b.overlap_with_zero()


def test_overlap_known_answers():
    """crates/ppvm-pauli-sum/src/sum/trace.rs:52-80 on the GPU."""
    from ppvm_amd import PauliSum

    def mk(terms):
        return PauliSum([t for t, _ in terms], 4, coefficients=[c for _, c in terms])

    a = mk([("IIII", 1.0), ("XIII", 2.0)])
    empty = mk([("XIII", 1e-20)])  # the constructor needs one term; truncation removes it
    empty.truncate()
    assert len(empty) == 0
    assert a.overlap(empty) == 0.0 and empty.overlap(a) == 0.0
    s = mk([("XIII", 3.0)])
    assert s.overlap(s) == 9.0
    assert mk([("XIII", 1.0)]).overlap(mk([("YIII", 1.0)])) == 0.0
    assert a.overlap(mk([("IIII", 3.0), ("XIII", 4.0)])) == 11.0


@pytest.mark.parametrize("n", [6, 70, 130, 300])
def test_overlap_between_sums(n):
    """Two sums propagated by the same truncated circuit from different
    (overlapping) starting sets: overlap(a, b), overlap(b, a), overlap(a, a)
    against the oracle (relative 1e-12: the GPU sums in another association)."""
    rng = np.random.default_rng(300 + n)
    strings = random_terms(rng, n, 120)
    prog = random_program(rng, n, 60, trunc_every=True)
    gs, rs = [], []
    for side in range(2):
        sel = strings[side * 40: side * 40 + 80]
        g, r = _pair(n, sel, rng.normal(size=len(sel)).tolist(), tau=1e-3)
        g.apply(prog)
        r.apply(prog)
        assert len(g) == len(r)
        gs.append(g)
        rs.append(r)
    for i, j in ((0, 1), (1, 0), (0, 0)):
        og, oo = gs[i].overlap(gs[j]), rs[i].overlap(rs[j])
        assert abs(og - oo) <= 1e-12 * max(1.0, abs(oo)), (i, j, og, oo)
    # after further gates on one side only
    gs[0].h(0)
    rs[0].apply(O.Program().add(O.H, 0).array())
    og, oo = gs[0].overlap(gs[1]), rs[0].overlap(rs[1])
    assert abs(og - oo) <= 1e-12 * max(1.0, abs(oo))


PATTERNS = ["Z?*", "X?*", "[XY]?*", "[XZ]*", "Z0", "Z?0Z?1X2", "[XY]?{2}Z2", "X{2}", "X{0}", "_*",
            "[XYZ]*", "Z?*X5", "Y3Z?*", "[YZ]?*Z?*", ""]


@pytest.mark.parametrize("n", [6, 70, 130])
def test_trace_patterns(n):
    """trace(pattern) on the GPU vs the oracle after a truncated random circuit
    (relative 1e-12: the GPU sums in another association)."""
    rng = np.random.default_rng(900 + n)
    strings = random_terms(rng, n, 60)
    # add low-support words so that anchored / star patterns have matches
    for q in range(min(n, 8)):
        strings.append("I" * q + "Z" + "I" * (n - q - 1))
        strings.append("I" * q + "X" + "I" * (n - q - 1))
    strings = list(dict.fromkeys(strings))
    g, r = _pair(n, strings, rng.normal(size=len(strings)).tolist(), tau=1e-4)
    prog = random_program(rng, n, 30, trunc_every=True, kinds=["h", "s", "cnot", "rx", "rz", "rzz"])
    g.apply(prog)
    r.apply(prog)
    assert len(g) == len(r)
    for pat in PATTERNS:
        tg, tr = g.trace(pat), r.trace(pat)
        assert abs(tg - tr) <= 1e-12 * max(1.0, abs(tr)), (pat, tg, tr)
    for bad in ("A", "[]", "X*Y"):
        with pytest.raises(ValueError):
            g.trace(bad)


@pytest.mark.parametrize("n,tau", [(5, 0.0), (5, 1e-3), (70, 1e-4), (130, 1e-4)])
def test_amplitude_damping(n, tau):
    """Amplitude damping mixed with rotations, vs the oracle after every op:
    identical keys, coefficient bits and order.  The start set holds I and Z
    versions of the same words so that damped Z terms merge into existing I terms."""
    rng = np.random.default_rng(500 + n)
    base = random_terms(rng, n, 40)
    strings = list(base)
    for w in base[:20]:
        q = int(rng.integers(0, n))
        strings.append(w[:q] + "Z" + w[q + 1:])
        strings.append(w[:q] + "I" + w[q + 1:])
    strings = list(dict.fromkeys(strings))
    g, r = _pair(n, strings, rng.normal(size=len(strings)).tolist(), tau=tau)
    p = O.Program()
    for i in range(24):
        q = int(rng.integers(0, n))
        if i % 3 == 0:
            p.add(O.AMPLITUDE_DAMPING, q, p=[float(rng.uniform(0.05, 0.6))])
        else:
            p.add([O.RX, O.RZ, O.H][i % 3], q, theta=float(rng.uniform(-2, 2)))
        if tau > 0:
            p.truncate()
    prog = p.array()
    marks = list(range(len(prog)))
    run_segments(g, r, prog, marks, ctx=f"damping n={n}")


def test_amplitude_damping_python_api():
    from ppvm_amd import PauliSum
    s = PauliSum(["ZI"], 2, coefficients=[1.0])
    s.amplitude_damping(0, 0.2)
    t = dict(s.terms)
    assert t["ZI"] == pytest.approx(0.8) and t["II"] == pytest.approx(0.2)
    s = PauliSum(["ZI"], 2, coefficients=[1.0])
    s.amplitude_damping(0, 0.0)
    assert dict(s.terms) == {"ZI": 1.0}
