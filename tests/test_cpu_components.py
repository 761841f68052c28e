"""CPU checks of host-side components: workload generators, sweep features,
and the MLP reference (pinned against finite differences)."""

import numpy as np

from ppvm_amd import ops as O
from ppvm_amd import workloads as WL


def test_workloads_are_deterministic_and_shaped():
    a, b = WL.lattice_tfim(), WL.lattice_tfim()
    assert np.array_equal(a.program, b.program)
    assert a.n_gates == 20 * (90 + 90 + 100)
    assert len(a.layer_marks) == 60
    # backward order: the first ops are rx on qubits 99..0, each followed by a truncate
    assert [int(c) for c in a.program["code"][:4]] == [O.RX, O.TRUNCATE, O.RX, O.TRUNCATE]
    assert [int(q) for q in a.program["q0"][:4:2]] == [99, 98]
    assert a.observable == [("I" * 55 + "Z" + "I" * 44, 1.0)]
    c = WL.chain_tfim()
    assert c.n_gates == 10 * (19 + 20) and c.observable[0][0][10] == "Z"


def test_heavy_hex_graph_and_random_circuit():
    n, edges = WL.heavy_hex_graph()
    assert n == 127 and len(edges) == 144
    deg = np.zeros(n, int)
    for a, b in edges:
        deg[a] += 1
        deg[b] += 1
    assert deg.max() <= 3 and deg.min() >= 1
    w = WL.random_circuit()
    w2 = WL.random_circuit()
    assert np.array_equal(w.program, w2.program)
    assert len(w.params["obs"]) == 3 and len(set(w.params["obs"])) == 3
    codes = set(int(c) for c in w.program["code"])
    assert codes <= {O.CNOT, O.H, O.S, O.RZ, O.TRUNCATE}


def test_batched_chain_parameters():
    ws = WL.batched_chains()
    assert len(ws) == 256
    Ls = [w.n_qubits for w in ws]
    assert min(Ls) >= 32 and max(Ls) <= 128
    steps = [w.params["steps"] for w in ws]
    assert min(steps) >= 5 and max(steps) <= 30
    hs = [w.params["h"] for w in ws]
    assert 0.5 <= min(hs) and max(hs) < 1.5


def test_sweep_features():
    from ppvm_amd.surrogate import features
    f = features([0.5, 4.0], [0.0, 2.0])
    assert f.shape == (2, 11)
    assert np.allclose(f[:, 0], [-1.0, 1.0])
    assert np.allclose(f[:, 2], [-1.0, 1.0])
    g = features([1.3, 1.3], [0.99, 1.01])  # both round to 20 steps: the same physical point
    assert np.array_equal(g[0], g[1])


def test_mlp_reference_matches_finite_differences():
    from oracle.mlp_ref import forward_backward, shapes
    rng = np.random.default_rng(0)
    hidden = 8
    K, N = shapes(hidden)
    P = sum(n * k + n for k, n in zip(K, N))
    p = rng.normal(scale=0.3, size=P)
    X = rng.uniform(-1, 1, size=(20, 5))
    Y = rng.normal(size=20)
    loss, g = forward_backward(p, X, Y, hidden, rounding=False)
    idx = rng.choice(P, size=40, replace=False)
    for i in idx:
        e = np.zeros(P)
        e[i] = 1e-6
        lp, _ = forward_backward(p + e, X, Y, hidden, rounding=False)
        lm, _ = forward_backward(p - e, X, Y, hidden, rounding=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    from oracle.mlp_ref import bf16
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 1e-40], dtype=np.float32)
    r = bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 4 * 2**-8) and r[3] == -2.5


def test_shipped_sweep_grid_matches_oracle_prefix():
    """The propagator grid shipped for the surrogate (written on the GPU by
    `python -m ppvm_amd.sweep`) agrees with the oracle over the first Trotter
    steps at the smallest and largest h (1e-10 absolute)."""
    from ppvm_amd import workloads as WL
    from ppvm_amd.surrogate import load_grid
    from tests.helpers import oracle_for
    d = load_grid()
    assert d["grid"].shape == (64, 41)
    assert np.all(d["grid"][:, 0] == 1.0)
    for i in (0, 63):
        w = WL.lattice_tfim(h=float(d["h"][i]), steps=9)
        o = oracle_for(w)
        vals, prev = [o.overlap_with_zero()], 0
        for li, m in enumerate(w.layer_marks):
            o.apply(w.program[prev:m + 1])
            prev = m + 1
            if li % 3 == 2:
                vals.append(o.overlap_with_zero())
        assert np.max(np.abs(np.array(vals) - d["grid"][i, :10])) <= 1e-10


def test_term_parser_forms():
    """PauliSum.new's term forms (ppvm-python/src/ppvm/paulisum.py:20-38):
    full strings pass through, the compact form names qubits by index (a later
    letter for the same qubit wins, unnamed qubits are I), tuples carry a
    coefficient, and an index past the register raises."""
    import pytest
    from ppvm_amd.paulisum import _parse_term
    assert _parse_term("XYZI", 4) == ("XYZI", 1.0)
    assert _parse_term("X3Z0", 5) == ("ZIIXI", 1.0)
    assert _parse_term(("Z10", 0.5), 12) == ("I" * 10 + "ZI", 0.5)
    assert _parse_term("X1Y1", 3) == ("IYI", 1.0)
    assert _parse_term("I0", 2) == ("II", 1.0)
    # not the compact form: kept as given (the constructor checks the length)
    assert _parse_term("X9Y", 3) == ("X9Y", 1.0)
    assert _parse_term("X", 3) == ("X", 1.0)
    with pytest.raises(ValueError, match="Qubit index 3 out of range for 3-qubit system."):
        _parse_term("Z3", 3)


def test_surrogate_split_is_disjoint_in_h_and_steps():
    """Training points are the seed-sampled ones; the validation pairs share no
    (h, steps) pair with the fitted points, and new query pairs none with
    either; known query pairs are exactly the query points on training pairs."""
    from ppvm_amd.surrogate import split_sweep
    from ppvm_amd.sweep import H_GRID, T_GRID
    sp = split_sweep(H_GRID, T_GRID, 2048, 31, 0.2)
    assert len(sp.train) == 2048 and len(sp.query) == 2048
    assert len(np.intersect1d(sp.train, sp.query)) == 0
    assert len(sp.fit) + len(sp.val) == 2048
    fit_p, val_p = set(sp.pair[sp.fit]), set(sp.pair[sp.val])
    assert not fit_p & val_p and abs(len(val_p) / len(fit_p | val_p) - 0.2) < 0.01
    new_p = set(sp.pair[sp.query[~sp.known]])
    assert not new_p & (fit_p | val_p)
    assert set(sp.pair[sp.query[sp.known]]) <= fit_p | val_p
    # pairs: h index and round(t / dt)
    i = sp.query[5]
    assert sp.steps[i] == round(T_GRID[i % 64] / 0.05) and sp.hidx[i] == i // 64
    # seed-31 draw of the configured sampler
    want = np.sort(np.random.default_rng(31).choice(np.arange(4096), size=2048, replace=False))
    assert np.array_equal(sp.train, want)


def test_surrogate_bound_calibration():
    from ppvm_amd.surrogate import calibrate_bound
    sd = np.array([0.5, 0.1, 0.3, 0.2, 0.4])
    err = np.array([0.0, 0.0, 2e-3, 0.0, 0.0])  # the third-most-confident point misses
    assert calibrate_bound(sd, err, 1e-3) == 0.2
    assert calibrate_bound(sd, np.zeros(5), 1e-3) == 0.5
    assert calibrate_bound(sd, np.array([0, 5e-3, 0, 0, 0.0]), 1e-3) == -1.0
    # every validation point under the bound is within the bar
    rng = np.random.default_rng(3)
    sd, err = rng.random(500), rng.random(500) * 2e-3
    b = calibrate_bound(sd, err, 1e-3)
    assert np.all(err[sd <= b] <= 1e-3) and np.any(err[sd > b] > 1e-3)
