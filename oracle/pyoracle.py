"""ctypes wrapper of the CPU oracle (``oracle/libppvm_oracle.so``).

TEST INFRASTRUCTURE ONLY: the parity oracle and the CPU baseline.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module, and only to check results against or as the timed CPU baseline.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from ppvm_amd.ops import (OP_DTYPE, pack_lossy_strings, pack_strings, unpack_lossy_words, unpack_words,
                          words_for)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libppvm_oracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the oracle with its Makefile (g++); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ppvo_create.restype = ctypes.c_void_p
        L.ppvo_create.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64]
        L.ppvo_destroy.argtypes = [ctypes.c_void_p]
        L.ppvo_clone.restype = ctypes.c_void_p
        L.ppvo_clone.argtypes = [ctypes.c_void_p]
        L.ppvo_words.restype = ctypes.c_uint32
        L.ppvo_words.argtypes = [ctypes.c_void_p]
        L.ppvo_set_preserve.argtypes = [ctypes.c_void_p, _u64p, _u64p, ctypes.c_size_t]
        L.ppvo_add_terms.argtypes = [ctypes.c_void_p, _u64p, _u64p, _f64p, ctypes.c_size_t]
        L.ppvo_apply.restype = ctypes.c_int
        L.ppvo_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.ppvo_overlap_with_zero.restype = ctypes.c_double
        L.ppvo_overlap_with_zero.argtypes = [ctypes.c_void_p]
        L.ppvo_overlap.restype = ctypes.c_double
        L.ppvo_overlap.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ppvo_trace.restype = ctypes.c_int
        L.ppvo_trace.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _f64p]
        L.ppvo_pattern_contains.restype = ctypes.c_int
        L.ppvo_pattern_contains.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _u64p, _u64p,
                                            ctypes.c_void_p, ctypes.c_size_t]
        L.ppvo_len.restype = ctypes.c_uint64
        L.ppvo_len.argtypes = [ctypes.c_void_p]
        L.ppvo_export.restype = ctypes.c_uint64
        L.ppvo_export.argtypes = [ctypes.c_void_p, _u64p, _u64p, _f64p, ctypes.c_size_t]
        L.ppvo_term_updates.restype = ctypes.c_uint64
        L.ppvo_term_updates.argtypes = [ctypes.c_void_p]
        L.ppvo_fingerprint.restype = None
        L.ppvo_fingerprint.argtypes = [ctypes.c_void_p, _u64p]
        L.ppvo_run_marked.restype = ctypes.c_size_t
        L.ppvo_run_marked.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _u64p,
                                      ctypes.c_size_t, _u64p, _f64p]
        L.ppvo_lossy_create.restype = ctypes.c_void_p
        L.ppvo_lossy_create.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64]
        L.ppvo_lossy_set_preserve.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, ctypes.c_size_t]
        L.ppvo_lossy_add_terms.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, _f64p, ctypes.c_size_t]
        L.ppvo_lossy_export.restype = ctypes.c_uint64
        L.ppvo_lossy_export.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, _f64p, ctypes.c_size_t]
        L.ppvo_apply_sharded.restype = ctypes.c_int
        L.ppvo_apply_sharded.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_int]
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(_u64p)


class OracleSum:
    """CPU oracle PauliSum: insertion-ordered map with the reference's semantics."""

    def __init__(self, n_qubits: int, min_abs_coeff: float = 0.0,
                 max_pauli_weight: int | None = None, _handle=None):
        self.n_qubits = n_qubits
        self.W = words_for(n_qubits)
        L = lib()
        mw = (1 << 64) - 1 if max_pauli_weight is None else int(max_pauli_weight)
        self._h = _handle if _handle is not None else L.ppvo_create(n_qubits, float(min_abs_coeff), mw)
        self.min_abs_coeff = min_abs_coeff
        self.max_pauli_weight = max_pauli_weight

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.ppvo_destroy(h)
            self._h = None

    def copy(self) -> "OracleSum":
        return OracleSum(self.n_qubits, self.min_abs_coeff, self.max_pauli_weight,
                         _handle=lib().ppvo_clone(self._h))

    def add_words(self, x, z, c):
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, self.W)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, self.W)
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        lib().ppvo_add_terms(self._h, _p64(x), _p64(z), c.ctypes.data_as(_f64p), len(c))

    def add_strings(self, strings, coeffs=None):
        x, z = pack_strings(strings, self.n_qubits)
        c = np.ones(len(strings)) if coeffs is None else np.asarray(coeffs, dtype=np.float64)
        self.add_words(x, z, c)

    def set_preserve(self, strings):
        x, z = pack_strings(strings, self.n_qubits)
        lib().ppvo_set_preserve(self._h, _p64(x), _p64(z), len(strings))

    def apply(self, program: np.ndarray):
        program = np.ascontiguousarray(program, dtype=OP_DTYPE)
        rc = lib().ppvo_apply(self._h, program.ctypes.data, len(program))
        if rc != 0:
            raise ValueError("oracle rejected an op (bad qubit index or op code)")

    def apply_sharded(self, program: np.ndarray, threads: int):
        program = np.ascontiguousarray(program, dtype=OP_DTYPE)
        rc = lib().ppvo_apply_sharded(self._h, program.ctypes.data, len(program), int(threads))
        if rc != 0:
            raise ValueError("oracle rejected an op")

    def run_marked(self, program: np.ndarray, marks):
        program = np.ascontiguousarray(program, dtype=OP_DTYPE)
        marks = np.ascontiguousarray(marks, dtype=np.uint64)
        out_len = np.zeros(len(marks), dtype=np.uint64)
        out_ovl = np.zeros(len(marks), dtype=np.float64)
        m = lib().ppvo_run_marked(self._h, program.ctypes.data, len(program), _p64(marks),
                                  len(marks), _p64(out_len), out_ovl.ctypes.data_as(_f64p))
        return out_len[:m], out_ovl[:m]

    def overlap_with_zero(self) -> float:
        return lib().ppvo_overlap_with_zero(self._h)

    def overlap(self, other: "OracleSum") -> float:
        return lib().ppvo_overlap(self._h, other._h)

    def trace(self, pattern: str) -> float:
        """Sum of the coefficients of terms matching a Pauli pattern
        (crates/ppvm-pauli-word/src/pattern/); ValueError if it does not parse."""
        out = np.zeros(1, dtype=np.float64)
        if lib().ppvo_trace(self._h, pattern.encode(), out.ctypes.data_as(_f64p)) != 0:
            raise ValueError(f"invalid Pauli pattern {pattern!r}")
        return float(out[0])

    def pattern_contains(self, pattern: str, strings) -> list:
        """Pattern membership of each Pauli string (on this sum's qubit count)."""
        x, z = pack_strings(list(strings), self.n_qubits)
        out = np.zeros(len(strings), dtype=np.uint8)
        if lib().ppvo_pattern_contains(self._h, pattern.encode(), _p64(x), _p64(z), out.ctypes.data,
                                       len(strings)) != 0:
            raise ValueError(f"invalid Pauli pattern {pattern!r}")
        return [bool(v) for v in out]

    def __len__(self) -> int:
        return int(lib().ppvo_len(self._h))

    def term_updates(self) -> int:
        return int(lib().ppvo_term_updates(self._h))

    def fingerprint(self) -> tuple:
        """(count, key digest sum, key+coefficient digest sum): an
        order-independent digest of the live terms."""
        out = np.zeros(3, dtype=np.uint64)
        lib().ppvo_fingerprint(self._h, _p64(out))
        return tuple(int(v) for v in out)

    def export(self):
        n = len(self)
        x = np.zeros((n, self.W), dtype=np.uint64)
        z = np.zeros((n, self.W), dtype=np.uint64)
        c = np.zeros(n, dtype=np.float64)
        if n:
            lib().ppvo_export(self._h, _p64(x), _p64(z), c.ctypes.data_as(_f64p), n)
        return x, z, c

    def terms(self):
        x, z, c = self.export()
        return list(zip(unpack_words(x, z, self.n_qubits), c.tolist()))


class LossyOracleSum(OracleSum):
    """CPU oracle of a PauliSum over lossy words (I X Y Z L)."""

    def __init__(self, n_qubits: int, min_abs_coeff: float = 0.0,
                 max_pauli_weight: int | None = None, max_loss_weight: int | None = None, _handle=None):
        mw = (1 << 64) - 1 if max_pauli_weight is None else int(max_pauli_weight)
        mlw = (1 << 64) - 1 if max_loss_weight is None else int(max_loss_weight)
        h = _handle if _handle is not None else lib().ppvo_lossy_create(n_qubits, float(min_abs_coeff), mw, mlw)
        super().__init__(n_qubits, min_abs_coeff, max_pauli_weight, _handle=h)
        self.max_loss_weight = max_loss_weight

    def copy(self) -> "LossyOracleSum":
        return LossyOracleSum(self.n_qubits, self.min_abs_coeff, self.max_pauli_weight,
                              self.max_loss_weight, _handle=lib().ppvo_clone(self._h))

    def add_lossy_words(self, x, z, l, c):
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, self.W)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, self.W)
        l = np.ascontiguousarray(l, dtype=np.uint64).reshape(-1, self.W)
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        lib().ppvo_lossy_add_terms(self._h, _p64(x), _p64(z), _p64(l), c.ctypes.data_as(_f64p), len(c))

    def add_strings(self, strings, coeffs=None):
        x, z, l = pack_lossy_strings(list(strings), self.n_qubits)
        c = np.ones(len(strings)) if coeffs is None else np.asarray(coeffs, dtype=np.float64)
        self.add_lossy_words(x, z, l, c)

    def set_preserve(self, strings):
        x, z, l = pack_lossy_strings(list(strings), self.n_qubits)
        lib().ppvo_lossy_set_preserve(self._h, _p64(x), _p64(z), _p64(l), len(strings))

    def export(self):
        n = len(self)
        x = np.zeros((n, self.W), dtype=np.uint64)
        z = np.zeros((n, self.W), dtype=np.uint64)
        l = np.zeros((n, self.W), dtype=np.uint64)
        c = np.zeros(n, dtype=np.float64)
        if n:
            lib().ppvo_lossy_export(self._h, _p64(x), _p64(z), _p64(l), c.ctypes.data_as(_f64p), n)
        return x, z, l, c

    def terms(self):
        x, z, l, c = self.export()
        return list(zip(unpack_lossy_words(x, z, l, self.n_qubits), c.tolist()))
