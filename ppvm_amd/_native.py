"""ctypes binding of ``libppvm_hip.so`` (the C ABI in ``include/ppvm_hip.h``).

The product path has no CPU fallback: if the HIP library is missing or no
gfx950 device is present, every call raises.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import ops as _ops
from .ops import OP_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libppvm_hip.so")

PPVM_OK, PPVM_EINVAL, PPVM_ENOMEM, PPVM_EHIP, PPVM_ESTATE, PPVM_ENODEV = range(6)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)
_H = ctypes.c_void_p


class PpvmStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "term_updates", "gates", "rotations", "n_entries", "n_live", "capacity",
        "peak_live", "syncs", "regrows", "compactions")]


class PpvmError(RuntimeError):
    """Error reported by the native library (code + message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"ppvm error {code}: {msg}")
        self.code = code


_lib = None


def _declare(L):
    def d(name, *args, res=ctypes.c_int):
        f = getattr(L, name)
        f.argtypes = list(args)
        f.restype = res

    d("ppvm_last_error", res=ctypes.c_char_p)
    d("ppvm_device_info", ctypes.c_char_p, ctypes.c_size_t)
    d("ppvm_create", ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
      ctypes.POINTER(_H))
    d("ppvm_destroy", _H)
    d("ppvm_clone", _H, ctypes.POINTER(_H))
    d("ppvm_words", _H, _u32p)
    d("ppvm_set_preserve", _H, _u64p, _u64p, ctypes.c_size_t)
    d("ppvm_add_terms", _H, _u64p, _u64p, _f64p, ctypes.c_size_t)
    d("ppvm_clear", _H)
    for g in ("x", "y", "z", "h", "s", "s_dag", "sqrt_x", "sqrt_x_dag", "sqrt_y", "sqrt_y_dag",
              "cnot", "cz", "cy"):
        d("ppvm_" + g, _H, _u32p, ctypes.c_size_t)
    for g in ("rx", "ry", "rz", "rxx", "ryy", "rzz"):
        d("ppvm_" + g, _H, _u32p, ctypes.c_size_t, ctypes.c_double)
    d("ppvm_r", _H, ctypes.c_uint32, ctypes.c_double, ctypes.c_double)
    d("ppvm_pauli_error", _H, ctypes.c_uint32, _f64p)
    d("ppvm_depolarize1", _H, ctypes.c_uint32, ctypes.c_double)
    d("ppvm_amplitude_damping", _H, ctypes.c_uint32, ctypes.c_double)
    d("ppvm_depolarize2", _H, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double)
    d("ppvm_two_qubit_pauli_error", _H, ctypes.c_uint32, ctypes.c_uint32, _f64p)
    d("ppvm_truncate", _H)
    d("ppvm_overlap_with_zero", _H, _f64p)
    d("ppvm_overlap", _H, _H, _f64p)
    d("ppvm_trace", _H, ctypes.c_char_p, _f64p)
    d("ppvm_len", _H, _u64p)
    d("ppvm_export_terms", _H, _u64p, _u64p, _f64p, ctypes.c_size_t,
      ctypes.POINTER(ctypes.c_size_t))
    d("ppvm_apply", _H, ctypes.c_void_p, ctypes.c_size_t)
    d("ppvm_synchronize", _H)
    d("ppvm_fingerprint", _H, _u64p)
    d("ppvm_get_stats", _H, ctypes.POINTER(PpvmStats))
    d("ppvm_get_counters", _H, _u64p)
    d("ppvm_profile_enable", _H, ctypes.c_int)
    d("ppvm_profile_read", _H, _f64p, _u64p)
    d("ppvm_lossy_last_error", res=ctypes.c_char_p)
    d("ppvm_lossy_create", ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
      ctypes.POINTER(_H))
    d("ppvm_lossy_destroy", _H)
    d("ppvm_lossy_clone", _H, ctypes.POINTER(_H))
    d("ppvm_lossy_words", _H, _u32p)
    d("ppvm_lossy_set_preserve", _H, _u64p, _u64p, _u64p, ctypes.c_size_t)
    d("ppvm_lossy_add_terms", _H, _u64p, _u64p, _u64p, _f64p, ctypes.c_size_t)
    d("ppvm_lossy_apply", _H, ctypes.c_void_p, ctypes.c_size_t)
    d("ppvm_lossy_len", _H, _u64p)
    d("ppvm_lossy_export_terms", _H, _u64p, _u64p, _u64p, _f64p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t))
    d("ppvm_lossy_overlap_with_zero", _H, _f64p)
    d("ppvm_lossy_trace", _H, ctypes.c_char_p, _f64p)
    d("ppvm_lossy_overlap", _H, _H, _f64p)


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(rc: int, err=None) -> None:
    if rc != PPVM_OK:
        msg = (err or lib().ppvm_last_error)().decode("utf-8", "replace")
        if rc == PPVM_EINVAL:
            raise ValueError(msg)
        if rc == PPVM_ENOMEM:
            raise MemoryError(msg)
        raise PpvmError(rc, msg)


def device_info() -> str:
    buf = ctypes.create_string_buffer(256)
    check(lib().ppvm_device_info(buf, 256))
    return buf.value.decode()


def u64p(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def u32p(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def f64p(a: np.ndarray):
    return a.ctypes.data_as(_f64p)


class Handle:
    """Owning wrapper of a ``ppvm_handle``."""

    def __init__(self, n_qubits: int, min_abs_coeff: float, max_pauli_weight: int | None,
                 capacity_hint: int = 0, _raw=None):
        L = lib()
        if _raw is not None:
            self.h = _raw
        else:
            h = _H()
            mw = (1 << 64) - 1 if max_pauli_weight is None else int(max_pauli_weight)
            check(L.ppvm_create(int(n_qubits), float(min_abs_coeff), mw, int(capacity_hint),
                                ctypes.byref(h)))
            self.h = h
        w = ctypes.c_uint32()
        check(L.ppvm_words(self.h, ctypes.byref(w)))
        self.W = int(w.value)

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.ppvm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def clone(self) -> "Handle":
        h = _H()
        check(lib().ppvm_clone(self.h, ctypes.byref(h)))
        return Handle(0, 0.0, None, _raw=h)

    def add_words(self, x: np.ndarray, z: np.ndarray, c: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, self.W)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, self.W)
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        if not (len(x) == len(z) == len(c)):
            raise ValueError("Initial terms and coefficients need to be of same length!")
        check(lib().ppvm_add_terms(self.h, u64p(x), u64p(z), f64p(c), len(c)))

    def clear(self):
        check(lib().ppvm_clear(self.h))

    def set_preserve(self, x: np.ndarray, z: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, self.W)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, self.W)
        check(lib().ppvm_set_preserve(self.h, u64p(x), u64p(z), len(x)))

    def gate(self, name: str, targets, theta: float | None = None):
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.int64).reshape(-1))
        if (t < 0).any():
            raise ValueError("negative qubit index")
        t = t.astype(np.uint32)
        f = getattr(lib(), "ppvm_" + name)
        n = len(t)
        if name in ("cnot", "cz", "cy", "rxx", "ryy", "rzz"):
            if n % 2:
                raise ValueError("Expected an even number of targets for a two-qubit gate")
            n //= 2
        if theta is None:
            check(f(self.h, u32p(t), n))
        else:
            check(f(self.h, u32p(t), n, float(theta)))

    def apply(self, program: np.ndarray):
        program = np.ascontiguousarray(program, dtype=OP_DTYPE)
        check(lib().ppvm_apply(self.h, program.ctypes.data, len(program)))

    def r(self, q: int, axis_angle: float, theta: float):
        check(lib().ppvm_r(self.h, int(q), float(axis_angle), float(theta)))

    def pauli_error(self, q: int, p):
        arr = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(3))
        check(lib().ppvm_pauli_error(self.h, int(q), f64p(arr)))

    def depolarize1(self, q: int, p: float):
        check(lib().ppvm_depolarize1(self.h, int(q), float(p)))

    def depolarize2(self, a: int, b: int, p: float):
        check(lib().ppvm_depolarize2(self.h, int(a), int(b), float(p)))

    def two_qubit_pauli_error(self, a: int, b: int, p):
        arr = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(15))
        check(lib().ppvm_two_qubit_pauli_error(self.h, int(a), int(b), f64p(arr)))

    def amplitude_damping(self, q: int, gamma: float):
        check(lib().ppvm_amplitude_damping(self.h, int(q), float(gamma)))

    def overlap(self, other: "Handle") -> float:
        return overlap(self, other)

    def truncate(self):
        check(lib().ppvm_truncate(self.h))

    def synchronize(self):
        check(lib().ppvm_synchronize(self.h))

    def overlap_with_zero(self) -> float:
        v = ctypes.c_double()
        check(lib().ppvm_overlap_with_zero(self.h, ctypes.byref(v)))
        return v.value

    def trace(self, pattern: str) -> float:
        v = ctypes.c_double()
        check(lib().ppvm_trace(self.h, pattern.encode(), ctypes.byref(v)))
        return v.value

    def __len__(self) -> int:
        v = ctypes.c_uint64()
        check(lib().ppvm_len(self.h, ctypes.byref(v)))
        return int(v.value)

    def export(self):
        n = len(self)
        x = np.zeros((n, self.W), dtype=np.uint64)
        z = np.zeros((n, self.W), dtype=np.uint64)
        c = np.zeros(n, dtype=np.float64)
        got = ctypes.c_size_t()
        check(lib().ppvm_export_terms(self.h, u64p(x), u64p(z), f64p(c), n, ctypes.byref(got)))
        return x, z, c

    def fingerprint(self) -> tuple:
        """(count, key digest sum, key+coefficient digest sum) of the live
        terms, computed on the device; order-independent."""
        out = np.zeros(3, dtype=np.uint64)
        check(lib().ppvm_fingerprint(self.h, u64p(out)))
        return tuple(int(v) for v in out)

    def stats(self) -> dict:
        s = PpvmStats()
        check(lib().ppvm_get_stats(self.h, ctypes.byref(s)))
        return {name: int(getattr(s, name)) for name, _ in PpvmStats._fields_}

    COUNTERS = ("term_updates", "gates", "rotations", "scanned", "wl_total", "pairs_total",
                "children_total", "kills_total", "map_entries", "map_launches", "scan_planes",
                "capacity", "syncs", "regrows", "compactions", "peak_live", "transposed", "pair_fallback",
                "small_runs", "small_steps")

    def counters(self) -> dict:
        out = np.zeros(len(self.COUNTERS), dtype=np.uint64)
        check(lib().ppvm_get_counters(self.h, u64p(out)))
        return {k: int(v) for k, v in zip(self.COUNTERS, out)}

    def profile(self, enable: bool):
        check(lib().ppvm_profile_enable(self.h, 1 if enable else 0))

    def profile_read(self):
        ms = np.zeros(8, dtype=np.float64)
        n = np.zeros(8, dtype=np.uint64)
        check(lib().ppvm_profile_read(self.h, f64p(ms), u64p(n)))
        return ms, n


def overlap(a: Handle, b: Handle) -> float:
    v = ctypes.c_double()
    check(lib().ppvm_overlap(a.h, b.h, ctypes.byref(v)))
    return v.value


def _lcheck(rc: int) -> None:
    check(rc, lambda: lib().ppvm_lossy_last_error())


class LossyHandle:
    """Owning wrapper of a ``ppvm_lossy_handle``: the same methods as
    :class:`Handle` (gates become ``ppvm_op`` programs) plus the loss channels."""

    def __init__(self, n_qubits: int, min_abs_coeff: float, max_pauli_weight: int | None,
                 max_loss_weight: int | None = None, capacity_hint: int = 0, _raw=None):
        L = lib()
        if _raw is not None:
            self.h = _raw
        else:
            h = _H()
            mw = (1 << 64) - 1 if max_pauli_weight is None else int(max_pauli_weight)
            mlw = (1 << 64) - 1 if max_loss_weight is None else int(max_loss_weight)
            _lcheck(L.ppvm_lossy_create(int(n_qubits), float(min_abs_coeff), mw, mlw, int(capacity_hint),
                                        ctypes.byref(h)))
            self.h = h
        w = ctypes.c_uint32()
        _lcheck(L.ppvm_lossy_words(self.h, ctypes.byref(w)))
        self.W = int(w.value)

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.ppvm_lossy_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def clone(self) -> "LossyHandle":
        h = _H()
        _lcheck(lib().ppvm_lossy_clone(self.h, ctypes.byref(h)))
        return LossyHandle(0, 0.0, None, _raw=h)

    def _words(self, *planes):
        return [np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, self.W) for a in planes]

    def add_words(self, x, z, l, c):
        x, z, l = self._words(x, z, l)
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        if not (len(x) == len(z) == len(l) == len(c)):
            raise ValueError("Initial terms and coefficients need to be of same length!")
        _lcheck(lib().ppvm_lossy_add_terms(self.h, u64p(x), u64p(z), u64p(l), f64p(c), len(c)))

    def set_preserve(self, x, z, l):
        x, z, l = self._words(x, z, l)
        _lcheck(lib().ppvm_lossy_set_preserve(self.h, u64p(x), u64p(z), u64p(l), len(x)))

    def apply(self, program: np.ndarray):
        program = np.ascontiguousarray(program, dtype=OP_DTYPE)
        _lcheck(lib().ppvm_lossy_apply(self.h, program.ctypes.data, len(program)))

    def _run(self, rows):
        prog = _ops.Program()
        for r in rows:
            prog.add(*r[:4], p=r[4] if len(r) > 4 else None)
        self.apply(prog.array())

    _CODES = {**_ops.CLIFFORD_1Q, **_ops.CLIFFORD_2Q, **_ops.ROTATION_1Q, **_ops.ROTATION_2Q}

    def gate(self, name: str, targets, theta: float | None = None):
        t = np.asarray(targets, dtype=np.int64).reshape(-1)
        if (t < 0).any():
            raise ValueError("negative qubit index")
        code = self._CODES[name]
        th = 0.0 if theta is None else float(theta)
        if code in _ops.TWO_QUBIT:
            if len(t) % 2:
                raise ValueError("Expected an even number of targets for a two-qubit gate")
            self._run([(code, int(a), int(b), th) for a, b in zip(t[0::2], t[1::2])])
        else:
            self._run([(code, int(q), 0, th) for q in t])

    def r(self, q: int, axis_angle: float, theta: float):
        """rz(phi); rx(theta); rz(-phi) (crates/ppvm-pauli-sum/src/sum/rot1.rs:135-143)."""
        self._run([(_ops.RZ, q, 0, axis_angle), (_ops.RX, q, 0, theta), (_ops.RZ, q, 0, -axis_angle)])

    def pauli_error(self, q: int, p):
        self._run([(_ops.PAULI_ERROR, q, 0, 0.0, list(np.asarray(p, dtype=np.float64).reshape(3)))])

    def depolarize1(self, q: int, p: float):
        self._run([(_ops.DEPOLARIZE1, q, 0, 0.0, [p])])

    def depolarize2(self, a: int, b: int, p: float):
        self._run([(_ops.DEPOLARIZE2, a, b, 0.0, [p])])

    def two_qubit_pauli_error(self, a: int, b: int, p):
        self._run([(_ops.PAULI_ERROR2, a, b, 0.0, list(np.asarray(p, dtype=np.float64).reshape(15)))])

    def amplitude_damping(self, q: int, gamma: float):
        self._run([(_ops.AMPLITUDE_DAMPING, q, 0, 0.0, [gamma])])

    def loss_channel(self, q: int, p: float):
        self._run([(_ops.LOSS, q, 0, 0.0, [p])])

    def correlated_loss_channel(self, a: int, b: int, p):
        self._run([(_ops.CORRELATED_LOSS, a, b, 0.0, list(np.asarray(p, dtype=np.float64).reshape(3)))])

    def reset_loss_channel(self, q: int):
        self._run([(_ops.RESET_LOSS, q, 0, 0.0)])

    def truncate(self):
        self._run([(_ops.TRUNCATE, 0, 0, 0.0)])

    def synchronize(self):
        pass

    def overlap_with_zero(self) -> float:
        v = ctypes.c_double()
        _lcheck(lib().ppvm_lossy_overlap_with_zero(self.h, ctypes.byref(v)))
        return v.value

    def trace(self, pattern: str) -> float:
        v = ctypes.c_double()
        _lcheck(lib().ppvm_lossy_trace(self.h, pattern.encode(), ctypes.byref(v)))
        return v.value

    def overlap(self, other: "LossyHandle") -> float:
        v = ctypes.c_double()
        _lcheck(lib().ppvm_lossy_overlap(self.h, other.h, ctypes.byref(v)))
        return v.value

    def __len__(self) -> int:
        v = ctypes.c_uint64()
        _lcheck(lib().ppvm_lossy_len(self.h, ctypes.byref(v)))
        return int(v.value)

    def export(self):
        n = len(self)
        x, z, l = (np.zeros((n, self.W), dtype=np.uint64) for _ in range(3))
        c = np.zeros(n, dtype=np.float64)
        got = ctypes.c_size_t()
        _lcheck(lib().ppvm_lossy_export_terms(self.h, u64p(x), u64p(z), u64p(l), f64p(c), n, ctypes.byref(got)))
        return x, z, l, c

    def stats(self) -> dict:
        return {"n_live": len(self)}
