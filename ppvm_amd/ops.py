"""Op records and Pauli-word packing shared by the host API, the workloads and
the tests.

An op is the ``ppvm_op`` record of ``include/ppvm_hip.h`` (144 bytes).  A
program is a numpy structured array of ops; it is what ``ppvm_apply`` consumes.

Pauli words cross the boundary as two ``(count, W)`` uint64 arrays, x and z,
with qubit ``i`` at bit ``i % 64`` of word ``i // 64`` — the Lsb0 layout of
the reference's ``PauliWord<[u8;N]>`` planes
(crates/ppvm-pauli-word/src/word/data.rs:43-68).  String form: character ``i``
is qubit ``i``; ``'_'`` is skipped (word/data.rs:334-377).
"""

from __future__ import annotations

import math

import numpy as np

# Op codes (values fixed by include/ppvm_hip.h).
NOP = 0
X, Y, Z, H, S, S_DAG, SQRT_X, SQRT_X_DAG, SQRT_Y, SQRT_Y_DAG = range(1, 11)
CNOT, CZ, CY = 11, 12, 13
RX, RY, RZ, RXX, RYY, RZZ = 20, 21, 22, 23, 24, 25
PAULI_ERROR, DEPOLARIZE1, DEPOLARIZE2, PAULI_ERROR2 = 40, 41, 42, 43
AMPLITUDE_DAMPING = 44
LOSS, CORRELATED_LOSS, RESET_LOSS = 45, 46, 47  # lossy sums only
TRUNCATE = 60

CLIFFORD_1Q = {
    "x": X, "y": Y, "z": Z, "h": H, "s": S, "s_dag": S_DAG, "sqrt_x": SQRT_X,
    "sqrt_x_dag": SQRT_X_DAG, "sqrt_y": SQRT_Y, "sqrt_y_dag": SQRT_Y_DAG,
}
CLIFFORD_2Q = {"cnot": CNOT, "cz": CZ, "cy": CY}
ROTATION_1Q = {"rx": RX, "ry": RY, "rz": RZ}
ROTATION_2Q = {"rxx": RXX, "ryy": RYY, "rzz": RZZ}
ROTATIONS = set(ROTATION_1Q.values()) | set(ROTATION_2Q.values())
TWO_QUBIT = (set(CLIFFORD_2Q.values()) | set(ROTATION_2Q.values())
             | {DEPOLARIZE2, PAULI_ERROR2, CORRELATED_LOSS})

OP_DTYPE = np.dtype(
    [
        ("code", "<u4"),
        ("q0", "<u4"),
        ("q1", "<u4"),
        ("reserved", "<u4"),
        ("theta", "<f8"),
        ("p", "<f8", (15,)),
    ],
    align=True,
)
assert OP_DTYPE.itemsize == 144


def words_for(n_qubits: int) -> int:
    """u64 words per plane for ``n_qubits`` (at least one)."""
    return max(1, (n_qubits + 63) // 64)


class Program:
    """Append-only list of ops, materialised as an ``OP_DTYPE`` array."""

    def __init__(self) -> None:
        self._rows: list[tuple] = []

    def __len__(self) -> int:
        return len(self._rows)

    def add(self, code: int, q0: int = 0, q1: int = 0, theta: float = 0.0, p=None) -> "Program":
        pp = [0.0] * 15
        if p is not None:
            for i, v in enumerate(p):
                pp[i] = float(v)
        self._rows.append((int(code), int(q0), int(q1), 0, float(theta), pp))
        return self

    def truncate(self) -> "Program":
        return self.add(TRUNCATE)

    def array(self) -> np.ndarray:
        a = np.zeros(len(self._rows), dtype=OP_DTYPE)
        for i, (c, q0, q1, r, th, p) in enumerate(self._rows):
            a[i]["code"] = c
            a[i]["q0"] = q0
            a[i]["q1"] = q1
            a[i]["theta"] = th
            a[i]["p"] = p
        return a


def pack_strings(strings, n_qubits: int):
    """Pack Pauli strings into ``(x, z)`` uint64 arrays of shape (count, W)."""
    W = words_for(n_qubits)
    cnt = len(strings)
    x = np.zeros((cnt, W), dtype=np.uint64)
    z = np.zeros((cnt, W), dtype=np.uint64)
    for t, s in enumerate(strings):
        q = 0
        for ch in s:
            if ch == "_":
                continue
            if q >= n_qubits:
                raise ValueError(f"Pauli string {s!r} longer than n_qubits={n_qubits}")
            w, b = divmod(q, 64)
            bit = np.uint64(1) << np.uint64(b)
            if ch == "X":
                x[t, w] |= bit
            elif ch == "Z":
                z[t, w] |= bit
            elif ch == "Y":
                x[t, w] |= bit
                z[t, w] |= bit
            elif ch != "I":
                raise ValueError(f"Invalid Pauli character: {ch}")
            q += 1
    return x, z


def unpack_words(x: np.ndarray, z: np.ndarray, n_qubits: int) -> list[str]:
    """Inverse of :func:`pack_strings` (one string of length n_qubits per row)."""
    out = []
    xs = np.asarray(x, dtype=np.uint64)
    zs = np.asarray(z, dtype=np.uint64)
    for t in range(xs.shape[0]):
        chars = []
        for q in range(n_qubits):
            w, b = divmod(q, 64)
            xb = (int(xs[t, w]) >> b) & 1
            zb = (int(zs[t, w]) >> b) & 1
            chars.append("IXZY"[xb | (zb << 1)])
        out.append("".join(chars))
    return out


def pack_lossy_strings(strings, n_qubits: int):
    """Pack lossy Pauli strings (I X Y Z L, '_' skipped; loss/data.rs:33-45)
    into ``(x, z, l)`` uint64 arrays of shape (count, W)."""
    W = words_for(n_qubits)
    cnt = len(strings)
    l = np.zeros((cnt, W), dtype=np.uint64)
    plain = []
    for t, s in enumerate(strings):
        chars = []
        q = 0
        for ch in s:
            if ch == "_":
                continue
            if ch == "L":
                if q >= n_qubits:
                    raise ValueError(f"Pauli string {s!r} longer than n_qubits={n_qubits}")
                w, b = divmod(q, 64)
                l[t, w] |= np.uint64(1) << np.uint64(b)
                ch = "I"
            chars.append(ch)
            q += 1
        plain.append("".join(chars))
    x, z = pack_strings(plain, n_qubits)
    return x, z, l


def unpack_lossy_words(x: np.ndarray, z: np.ndarray, l: np.ndarray, n_qubits: int) -> list[str]:
    """Inverse of :func:`pack_lossy_strings`."""
    out = unpack_words(x, z, n_qubits)
    ls = np.asarray(l, dtype=np.uint64)
    for t in range(ls.shape[0]):
        if not ls[t].any():
            continue
        chars = list(out[t])
        for q in range(n_qubits):
            w, b = divmod(q, 64)
            if (int(ls[t, w]) >> b) & 1:
                chars[q] = "L"
        out[t] = "".join(chars)
    return out


def weight_of(x: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Pauli weight popcount(x|z) per row."""
    v = np.bitwise_or(np.asarray(x, dtype=np.uint64), np.asarray(z, dtype=np.uint64))
    return np.bitwise_count(v).sum(axis=1).astype(np.int64)


def storage_tier(n_qubits: int) -> int:
    """Storage tier k of the reference binding (smallest k with 2**k > ceil(n/8)
    bytes, ppvm-python/src/ppvm/paulisum.py:176-182)."""
    nbytes = math.ceil(n_qubits / 8.0)
    return next(k for k in range(16) if 2**k > nbytes)
