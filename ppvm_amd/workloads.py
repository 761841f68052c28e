"""Synthetic circuits of the measurement configurations, as op programs.

Every workload is propagated backwards (Heisenberg picture): the forward
circuit's gates are emitted in reverse order, with a TRUNCATE after every gate
(the per-gate cadence of crates/ppvm-pauli-sum/benches/trotter.rs:27-44 and
tests/trotter.rs:35-54).  Random choices come from a project-owned splitmix64
generator so every circuit is reproducible bit for bit.

Configurations (sizes as measured):
  chain_tfim     1D open-chain TFIM, n=20, J=1, h=1, dt=0.1, 10 steps, Z_10,
                 threshold 1e-4, no weight cap.
  lattice_tfim   10x10 open square-lattice TFIM (qubit 10*r+c), J=1, h=3.04,
                 dt=0.05, 20 steps, Z at (5,5) = qubit 55, threshold 1e-5,
                 max weight 10.
  random_circuit 127 qubits on a degree<=3 heavy-hex-like graph (seed 7),
                 30 layers of {maximal CNOT matching, H-or-S per qubit, rz on
                 Bernoulli(0.3) qubits} (seed 11), random weight-3 Z
                 observable (seed 13), threshold 1e-4, max weight 12.
  batched_chains 256 chains, L~U{32..128}, h~U[0.5,1.5], steps~U{5..30}
                 (seed 21), J=1, dt=0.1, Z at L//2, threshold 1e-5.
  lossy_lattice_tfim  the lattice TFIM on lossy words, 4 steps, loss 0.01 on
                 every qubit after each step, reset-loss at readout, max loss
                 weight 2.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import ops as O


class SplitMix64:
    """splitmix64 stream; ``uniform()`` is in [0, 1) with 53 random bits."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def shuffle(self, items: list) -> None:
        for i in range(len(items) - 1, 0, -1):
            j = int(self.uniform() * (i + 1))
            items[i], items[j] = items[j], items[i]


@dataclass
class Workload:
    name: str
    n_qubits: int
    min_abs_coeff: float
    max_pauli_weight: int | None
    observable: list[tuple[str, float]]
    program: np.ndarray
    layer_marks: list[int] = field(default_factory=list)  # op index ending each layer
    params: dict = field(default_factory=dict)

    @property
    def n_gates(self) -> int:
        return int(np.count_nonzero(self.program["code"] != O.TRUNCATE))


def _z_string(n: int, qubits) -> str:
    s = ["I"] * n
    for q in qubits:
        s[q] = "Z"
    return "".join(s)


def _emit_backward(forward: list[tuple], prog: O.Program, layer_ends: set[int], marks: list[int]):
    """Append the reversed forward gate list, truncating after every gate.
    ``layer_ends`` holds forward indices that START a forward layer; in the
    reversed order the mark is placed after the gate with that index."""
    for fi in range(len(forward) - 1, -1, -1):
        code, q0, q1, th = forward[fi]
        prog.add(code, q0, q1, theta=th)
        prog.truncate()
        if fi in layer_ends:
            marks.append(len(prog) - 1)


def chain_tfim(n: int = 20, J: float = 1.0, h: float = 1.0, dt: float = 0.1, steps: int = 10,
               obs: int | None = None, min_abs_coeff: float = 1e-4,
               max_pauli_weight: int | None = None) -> Workload:
    """1D open-chain TFIM: forward step = rzz(2 J dt) on (i,i+1), then rx(2 h dt)."""
    obs = n // 2 if obs is None else obs
    fwd: list[tuple] = []
    starts: set[int] = set()
    for _ in range(steps):
        starts.add(len(fwd))
        for i in range(n - 1):
            fwd.append((O.RZZ, i, i + 1, 2.0 * J * dt))
        starts.add(len(fwd))
        for i in range(n):
            fwd.append((O.RX, i, 0, 2.0 * h * dt))
    prog = O.Program()
    marks: list[int] = []
    _emit_backward(fwd, prog, starts, marks)
    return Workload("chain_tfim", n, min_abs_coeff, max_pauli_weight, [(_z_string(n, [obs]), 1.0)],
                    prog.array(), marks,
                    dict(n=n, J=J, h=h, dt=dt, steps=steps, obs=obs))


def noisy_chain_tfim(n: int, J: float = 1.0, h: float = 1.0, steps: int = 20, dt: float = 0.1,
                     min_abs_coeff: float = 1e-6, noise: float = 1e-4) -> Workload:
    """The circuit of crates/ppvm-pauli-sum/examples/trotter_storage_cliff.rs:44-95
    (the reference's only recorded timings): observable sum_i Z_i; per step,
    for every qubit pauli_error(i, noise/4 each) then rx(i, h dt), then for
    every bond pauli_error(i+1), pauli_error(i), rzz(i, i+1, J dt); truncation
    (threshold only) after every call.  The calls are applied as written, in
    the reference's order."""
    p3 = [noise / 4.0] * 3
    prog = O.Program()
    marks: list[int] = []
    for _ in range(steps):
        for i in range(n):
            prog.add(O.PAULI_ERROR, i, p=p3).truncate()
            prog.add(O.RX, i, theta=dt * h).truncate()
        for i in range(n - 1):
            prog.add(O.PAULI_ERROR, i + 1, p=p3).truncate()
            prog.add(O.PAULI_ERROR, i, p=p3).truncate()
            prog.add(O.RZZ, i, i + 1, theta=dt * J).truncate()
        marks.append(len(prog) - 1)
    obs = [(_z_string(n, [i]), 1.0) for i in range(n)]
    return Workload("noisy_chain_tfim", n, min_abs_coeff, None, obs, prog.array(), marks,
                    dict(n=n, J=J, h=h, steps=steps, dt=dt, noise=noise))


def lattice_bonds(L: int):
    horiz = [(L * r + c, L * r + c + 1) for r in range(L) for c in range(L - 1)]
    vert = [(L * r + c, L * (r + 1) + c) for r in range(L - 1) for c in range(L)]
    return horiz, vert


def lattice_tfim(L: int = 10, J: float = 1.0, h: float = 3.04, dt: float = 0.05, steps: int = 20,
                 obs: tuple[int, int] = (5, 5), min_abs_coeff: float = 1e-5,
                 max_pauli_weight: int | None = 10) -> Workload:
    """2D open square-lattice TFIM: forward step = rzz on horizontal bonds,
    rzz on vertical bonds, rx on all qubits."""
    n = L * L
    horiz, vert = lattice_bonds(L)
    fwd: list[tuple] = []
    starts: set[int] = set()
    for _ in range(steps):
        starts.add(len(fwd))
        for a, b in horiz:
            fwd.append((O.RZZ, a, b, 2.0 * J * dt))
        starts.add(len(fwd))
        for a, b in vert:
            fwd.append((O.RZZ, a, b, 2.0 * J * dt))
        starts.add(len(fwd))
        for q in range(n):
            fwd.append((O.RX, q, 0, 2.0 * h * dt))
    prog = O.Program()
    marks: list[int] = []
    _emit_backward(fwd, prog, starts, marks)
    oq = L * obs[0] + obs[1]
    return Workload("lattice_tfim", n, min_abs_coeff, max_pauli_weight, [(_z_string(n, [oq]), 1.0)],
                    prog.array(), marks,
                    dict(L=L, J=J, h=h, dt=dt, steps=steps, obs_qubit=oq))


def lossy_lattice_tfim(L: int = 10, J: float = 1.0, h: float = 3.04, dt: float = 0.05, steps: int = 4,
                       p_loss: float = 0.01, obs: tuple[int, int] = (5, 5), min_abs_coeff: float = 1e-5,
                       max_pauli_weight: int | None = 10, max_loss_weight: int = 2) -> Workload:
    """The lattice TFIM on lossy words: each forward step is followed by a loss
    channel (probability p_loss) on every qubit, and the readout resets lost
    qubits to |0> (reset-loss on every qubit at the end of the circuit, i.e.
    first in the backward program).  Terms with more than max_loss_weight lost
    qubits are truncated."""
    n = L * L
    horiz, vert = lattice_bonds(L)
    fwd: list[tuple] = []
    starts: set[int] = set()
    for _ in range(steps):
        starts.add(len(fwd))
        fwd += [(O.RZZ, a, b, 2.0 * J * dt, None) for a, b in horiz + vert]
        fwd += [(O.RX, q, 0, 2.0 * h * dt, None) for q in range(n)]
        fwd += [(O.LOSS, q, 0, 0.0, [p_loss]) for q in range(n)]
    fwd += [(O.RESET_LOSS, q, 0, 0.0, None) for q in range(n)]
    prog = O.Program()
    marks: list[int] = []
    for fi in range(len(fwd) - 1, -1, -1):
        code, q0, q1, th, pp = fwd[fi]
        prog.add(code, q0, q1, theta=th, p=pp)
        prog.truncate()
        if fi in starts:
            marks.append(len(prog) - 1)
    oq = L * obs[0] + obs[1]
    return Workload("lossy_lattice_tfim", n, min_abs_coeff, max_pauli_weight, [(_z_string(n, [oq]), 1.0)],
                    prog.array(), marks,
                    dict(L=L, J=J, h=h, dt=dt, steps=steps, p_loss=p_loss, obs_qubit=oq,
                         max_loss_weight=max_loss_weight))


def heavy_hex_graph(seed: int = 7) -> tuple[int, list[tuple[int, int]]]:
    """127-qubit heavy-hex-like coupling graph (max degree 3): 7 rows of
    14/15/15/15/15/15/14 qubits joined by 4 bridge qubits per row gap (columns
    0,4,8,12 below even rows, 2,6,10,14 below odd rows), with the qubit labels
    permuted by a splitmix64 shuffle with ``seed``."""
    rows = [14, 15, 15, 15, 15, 15, 14]
    label = 0
    row_ids: list[dict[int, int]] = []
    edges: list[tuple[int, int]] = []
    for r, ln in enumerate(rows):
        start_col = 1 if r == len(rows) - 1 else 0
        ids = {}
        for c in range(start_col, start_col + ln):
            ids[c] = label
            label += 1
        row_ids.append(ids)
        cols = sorted(ids)
        for c0, c1 in zip(cols[:-1], cols[1:]):
            edges.append((ids[c0], ids[c1]))
    for g in range(len(rows) - 1):
        off = 0 if g % 2 == 0 else 2
        for k in range(4):
            col = 4 * k + off
            b = label
            label += 1
            edges.append((row_ids[g][col], b))
            edges.append((b, row_ids[g + 1][col]))
    n = label
    perm = list(range(n))
    SplitMix64(seed).shuffle(perm)
    edges = [(perm[a], perm[b]) for a, b in edges]
    deg = [0] * n
    for a, b in edges:
        deg[a] += 1
        deg[b] += 1
    assert n == 127 and max(deg) <= 3, (n, max(deg))
    return n, edges


def random_circuit(layers: int = 30, graph_seed: int = 7, gate_seed: int = 11, obs_seed: int = 13,
                   min_abs_coeff: float = 1e-4, max_pauli_weight: int | None = 12,
                   rz_prob: float = 0.3) -> Workload:
    """Clifford + rz circuit on the heavy-hex-like graph.  Per layer: a random
    maximal matching of CNOTs (random orientation), then exactly one of H / S
    (probability 0.5 each) on every qubit, then rz(theta ~ U[0, 2 pi)) on each
    qubit independently with probability ``rz_prob``."""
    n, edges = heavy_hex_graph(graph_seed)
    rng = SplitMix64(gate_seed)
    fwd: list[tuple] = []
    starts: set[int] = set()
    for _ in range(layers):
        starts.add(len(fwd))
        order = list(edges)
        rng.shuffle(order)
        used = [False] * n
        for a, b in order:
            if used[a] or used[b]:
                continue
            used[a] = used[b] = True
            if rng.uniform() < 0.5:
                fwd.append((O.CNOT, a, b, 0.0))
            else:
                fwd.append((O.CNOT, b, a, 0.0))
        for q in range(n):
            fwd.append((O.H if rng.uniform() < 0.5 else O.S, q, 0, 0.0))
        for q in range(n):
            if rng.uniform() < rz_prob:
                fwd.append((O.RZ, q, 0, 2.0 * math.pi * rng.uniform()))
    orng = SplitMix64(obs_seed)
    qs: list[int] = []
    while len(qs) < 3:
        q = orng.randint(0, n - 1)
        if q not in qs:
            qs.append(q)
    prog = O.Program()
    marks: list[int] = []
    _emit_backward(fwd, prog, starts, marks)
    return Workload("random_circuit", n, min_abs_coeff, max_pauli_weight, [(_z_string(n, qs), 1.0)],
                    prog.array(), marks, dict(layers=layers, obs=qs, edges=len(edges)))


def batched_chains(count: int = 256, seed: int = 21, J: float = 1.0, dt: float = 0.1,
                   min_abs_coeff: float = 1e-5) -> list[Workload]:
    """Independent 1D TFIM instances with L~U{32..128}, h~U[0.5,1.5],
    steps~U{5..30}, observable Z at L//2."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        L = rng.randint(32, 128)
        h = 0.5 + rng.uniform()
        steps = rng.randint(5, 30)
        w = chain_tfim(n=L, J=J, h=h, dt=dt, steps=steps, obs=L // 2, min_abs_coeff=min_abs_coeff,
                       max_pauli_weight=None)
        w.name = f"batched_chain_{i}"
        out.append(w)
    return out


def truncate_steps(w: Workload, steps: int) -> Workload:
    """The first ``steps`` backward Trotter steps of a TFIM workload (a prefix
    of the program; used to time identical prefixes on CPU and GPU)."""
    per = 3 if w.name == "lattice_tfim" else 2
    marks = w.layer_marks[: per * steps]
    end = marks[-1] + 1 if marks else 0
    return Workload(w.name + f"_prefix{steps}", w.n_qubits, w.min_abs_coeff, w.max_pauli_weight,
                    w.observable, w.program[:end].copy(), marks, dict(w.params, steps=steps))
