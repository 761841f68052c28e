"""Shared helpers for the parity tests (GPU engine vs CPU oracle)."""

from __future__ import annotations

import functools

import numpy as np

from oracle.pyoracle import OracleSum
from ppvm_amd import ops as O


def oracle_for(w) -> OracleSum:
    s = OracleSum(w.n_qubits, w.min_abs_coeff, w.max_pauli_weight)
    s.add_strings([t for t, _ in w.observable], [c for _, c in w.observable])
    return s


def gpu_for(w, capacity_hint: int = 0):
    from ppvm_amd import PauliSum
    return PauliSum.new(w.n_qubits, w.observable, min_abs_coeff=w.min_abs_coeff,
                        max_pauli_weight=w.max_pauli_weight, capacity_hint=capacity_hint)


def assert_same_terms(gpu_export, oracle_export, ordered: bool = True, ctx: str = ""):
    """Bit-identical keys and coefficients (and, if ordered, identical order)."""
    gx, gz, gc = gpu_export
    ox, oz, oc = oracle_export
    assert len(gc) == len(oc), f"{ctx}: {len(gc)} GPU terms vs {len(oc)} oracle terms"
    if not ordered:
        def keyed(x, z, c):
            k = np.concatenate([x, z], axis=1)
            order = np.lexsort(k.T[::-1])
            return k[order], c[order]
        gk, gc = keyed(gx, gz, gc)
        ok, oc = keyed(ox, oz, oc)
        assert np.array_equal(gk, ok), f"{ctx}: term sets differ"
    else:
        assert np.array_equal(gx, ox) and np.array_equal(gz, oz), f"{ctx}: keys/order differ"
    gb = gc.view(np.uint64)
    ob = oc.view(np.uint64)
    bad = np.nonzero(gb != ob)[0]
    assert len(bad) == 0, f"{ctx}: {len(bad)} coefficients differ, first {gc[bad[0]]!r} vs {oc[bad[0]]!r}"


def run_segments(gpu, ref: OracleSum, program: np.ndarray, marks, check_terms=True, ordered=True,
                 ctx=""):
    """Apply program up to each mark on both sides and compare."""
    prev = 0
    for li, m in enumerate(marks):
        seg = program[prev:m + 1]
        gpu.apply(seg)
        ref.apply(seg)
        prev = m + 1
        assert len(gpu) == len(ref), f"{ctx} layer {li}: len {len(gpu)} vs {len(ref)}"
        og, oo = gpu.overlap_with_zero(), ref.overlap_with_zero()
        assert abs(og - oo) <= 1e-10, f"{ctx} layer {li}: overlap {og!r} vs {oo!r}"
        if check_terms:
            assert_same_terms(gpu._export(), ref.export(), ordered, f"{ctx} layer {li}")
    if prev < len(program):
        gpu.apply(program[prev:])
        ref.apply(program[prev:])


def random_terms(rng, n, count):
    """count distinct random Pauli strings on n qubits."""
    seen = set()
    out = []
    while len(out) < count:
        s = "".join("IXYZ"[v] for v in rng.integers(0, 4, size=n))
        if s not in seen:
            seen.add(s)
            out.append(s)
    return out


def random_program(rng, n, ngates, trunc_every=True, kinds=None):
    p = O.Program()
    kinds = kinds or ["x", "y", "z", "h", "s", "s_dag", "sqrt_x", "sqrt_x_dag", "sqrt_y",
                      "sqrt_y_dag", "cnot", "cz", "cy", "rx", "ry", "rz", "rxx", "ryy", "rzz"]
    codes = {**O.CLIFFORD_1Q, **O.CLIFFORD_2Q, **O.ROTATION_1Q, **O.ROTATION_2Q}
    for _ in range(ngates):
        k = kinds[rng.integers(0, len(kinds))]
        c = codes[k]
        a = int(rng.integers(0, n))
        b = int(rng.integers(0, n - 1))
        if b >= a:
            b += 1
        th = float(rng.uniform(-np.pi, np.pi))
        p.add(c, a, b if c in O.TWO_QUBIT else 0, theta=th)
        if trunc_every:
            p.truncate()
    return p.array()


def fingerprint(export) -> dict:
    """Order-independent fingerprint of a term set: count, SHA-256 of the keys
    sorted lexicographically (x words, then z words) and of the coefficient
    bits in that order."""
    import hashlib
    x, z, c = export
    x = np.asarray(x, dtype=np.uint64)
    z = np.asarray(z, dtype=np.uint64)
    if x.ndim == 1:  # flat term-major words
        x = x.reshape(len(c), -1) if len(c) else x.reshape(0, 1)
        z = z.reshape(len(c), -1) if len(c) else z.reshape(0, 1)
    k = np.concatenate([x, z], axis=1)
    order = np.lexsort(k.T[::-1]) if len(c) else np.zeros(0, dtype=np.int64)
    ks = np.ascontiguousarray(k[order])
    cs = np.ascontiguousarray(np.asarray(c, dtype=np.float64)[order])
    return {"terms": int(len(c)),
            "keys": hashlib.sha256(ks.tobytes()).hexdigest()[:32],
            "coefs": hashlib.sha256(cs.view(np.uint64).tobytes()).hexdigest()[:32]}


@functools.lru_cache(maxsize=1)
def fingerprint_workloads():
    """The workloads whose per-layer fingerprints are committed as fixtures:
    name -> Workload."""
    from ppvm_amd import workloads as WL
    ws = {"chain_tfim_n20": WL.chain_tfim(),
          "lattice_tfim_10x10_prefix6": WL.truncate_steps(WL.lattice_tfim(), 6),
          "random_circuit_127q": WL.random_circuit()}
    chains = WL.batched_chains()
    for i in (0, 1, 2, 3):
        ws[f"batched_chain_{i}"] = chains[i]
    return ws


def layer_fingerprints(sum_, w, export_fn) -> list:
    """Apply w layer by layer on sum_ and fingerprint after every layer."""
    out, prev = [], 0
    for m in w.layer_marks:
        sum_.apply(w.program[prev:m + 1])
        prev = m + 1
        fp = fingerprint(export_fn())
        fp["overlap"] = float(sum_.overlap_with_zero())
        out.append(fp)
    return out


def host_fingerprint(x, z, c):
    """ppvm_fingerprint's definition in numpy: per term, d = m(...m(m(s ^ x_0) ^ x_1)... ^ z_{W-1})
    with m the SplitMix64 finaliser, e = m(d ^ coefficient bits); (count, sum d, sum e) mod 2^64."""
    def m(v):
        v = (v ^ (v >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        v = (v ^ (v >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return v ^ (v >> np.uint64(31))
    with np.errstate(over="ignore"):
        d = np.full(len(c), np.uint64(0x9E3779B97F4A7C15))
        for col in list(np.asarray(x, dtype=np.uint64).T) + list(np.asarray(z, dtype=np.uint64).T):
            d = m(d ^ col)
        e = m(d ^ np.ascontiguousarray(c, dtype=np.float64).view(np.uint64))
        return (len(c), int(d.sum(dtype=np.uint64)), int(e.sum(dtype=np.uint64)))
