"""GPU classes with the surface of the reference's native module ``ppvm._core``.

The reference's Python ``PauliSum`` picks a native class by name,
``getattr(_core, f"PauliSumIndexMapFxHash{k}")`` (or ``PauliSumLossIndexMapFxHash{k}``
for lossy words), where ``2**k`` is the smallest power of two above the
number of bytes per plane, ``ceil(n_qubits / 8)`` (ppvm-python/src/ppvm/paulisum.py:155-189).
This module defines those 32 names, with the constructor and methods of
``ppvm-python/src/ppvm/_core.pyi:3-58`` / ``crates/ppvm-python-native/src/interface.rs:70-470``:
targets as a list (two-qubit ops as a flat list of pairs), the angle or
probabilities next, and ``truncate=True`` after every call.  Installed as the
reference package's ``ppvm._core`` (``sys.modules["ppvm._core"] = ppvm_amd.core``
before ``import ppvm``), the reference wrapper runs on the GPU unchanged.

All tiers run on the same engine (the word count follows ``n_qubits``); a
tier accepts at most ``8 * 2**k`` qubits, the capacity of its storage.
Differences from the native module: a length mismatch of ``terms`` and
``coefficients`` raises ``ValueError`` where the native module panics
(interface.rs:100-104).
"""

from __future__ import annotations

import numpy as np

from . import _native
from .ops import pack_lossy_strings, pack_strings, unpack_lossy_words, unpack_words, weight_of
from .paulisum import _rust_fixed3

UMAX = (1 << 64) - 1
TIERS = range(16)


def _pairs(targets) -> list[int]:
    t = [int(v) for v in targets]
    if len(t) % 2:
        raise ValueError("two-qubit operations require an even number of targets")  # lib.rs:12-22
    return t


class _PauliSumBase:
    """``_PauliSumBase`` of _core.pyi:3-58 on the GPU engine."""

    _tier_bytes = 1
    _lossy = False

    def __init__(self, n_qubits: int, min_abs_coeff: float = 1e-10, max_pauli_weight: int = UMAX,
                 max_loss_weight: int = UMAX, terms=(), coefficients=(), preserve_strings=()):
        n = int(n_qubits)
        if n > 8 * self._tier_bytes:
            raise ValueError(f"{type(self).__name__} stores {8 * self._tier_bytes} qubits; got n_qubits={n}")
        terms, coefficients = list(terms), [float(c) for c in coefficients]
        if len(terms) != len(coefficients):
            raise ValueError(f"{len(terms)} terms but {len(coefficients)} coefficients")
        self.n_qubits = n
        mw = None if int(max_pauli_weight) >= UMAX else int(max_pauli_weight)
        self._h = self._handle(n, float(min_abs_coeff), mw, int(max_loss_weight))
        if preserve_strings:
            self._h.set_preserve(*self._pack(list(preserve_strings)))
        if terms:
            self._h.add_words(*self._pack(terms), np.asarray(coefficients, dtype=np.float64))

    # ---- storage ---------------------------------------------------------------------
    def _handle(self, n, tau, mw, mlw):
        return _native.Handle(n, tau, mw)

    def _pack(self, strings):
        return pack_strings(strings, self.n_qubits)

    def _words(self):
        x, z, c = self._h.export()
        return unpack_words(x, z, self.n_qubits), weight_of(x, z), c

    @classmethod
    def _from_handle(cls, n_qubits: int, h):
        new = object.__new__(cls)
        new.n_qubits = n_qubits
        new._h = h
        return new

    # ---- Python protocol --------------------------------------------------------------
    def __str__(self) -> str:
        # Display of PauliSum: "{:.3} * {word}" joined by " + ", stably sorted by
        # weight (crates/ppvm-pauli-sum/src/sum/display.rs:29-44)
        words, w, c = self._words()
        order = sorted(range(len(words)), key=lambda i: int(w[i]))
        return " + ".join(f"{_rust_fixed3(float(c[i]))} * {words[i]}" for i in order)

    def __repr__(self) -> str:
        return f"PauliSum({self})"  # interface.rs:112-115

    def __copy__(self):
        return self._from_handle(self.n_qubits, self._h.clone())

    def __deepcopy__(self, memo):
        return self.__copy__()

    def __len__(self) -> int:
        return len(self._h)

    def __eq__(self, other):
        # PartialEq of PauliSum: same qubit count, same key -> value map (order ignored)
        if not isinstance(other, _PauliSumBase):
            return NotImplemented
        return self.n_qubits == other.n_qubits and dict(self.terms()) == dict(other.terms())

    def __ne__(self, other):
        eq = self.__eq__(other)
        return eq if eq is NotImplemented else not eq

    def _cmp(self, other):
        raise NotImplementedError("Only equality and inequality comparisons are supported for PauliSum.")

    __lt__ = __le__ = __gt__ = __ge__ = _cmp
    __hash__ = None

    # ---- observables ------------------------------------------------------------------
    def trace(self, pattern: str) -> float:
        return self._h.trace(pattern)

    def overlap_with_zero(self) -> float:
        return self._h.overlap_with_zero()

    def overlap(self, other) -> float:
        if type(other._h) is not type(self._h):
            raise TypeError("overlap needs two sums of the same kind")
        return self._h.overlap(other._h)

    def truncate(self) -> None:
        self._h.truncate()

    def terms(self) -> list[tuple[str, float]]:
        words, _, c = self._words()
        return list(zip(words, c.tolist()))

    def weights(self) -> list[tuple[str, int]]:
        words, w, _ = self._words()
        return list(zip(words, [int(v) for v in w]))

    def current_max_weight(self) -> int:
        _, w, _ = self._words()
        return int(w.max()) if len(w) else 0

    # ---- gates and channels: targets, then the angle / probabilities, then truncate ----
    def _after(self, truncate: bool) -> None:
        if truncate:
            self._h.truncate()

    def _one(self, name, targets, theta, truncate):
        self._h.gate(name, [int(t) for t in targets], theta)
        self._after(truncate)

    def _two(self, name, targets, theta, truncate):
        self._h.gate(name, _pairs(targets), theta)
        self._after(truncate)

    def x(self, targets, truncate: bool = True): self._one("x", targets, None, truncate)
    def y(self, targets, truncate: bool = True): self._one("y", targets, None, truncate)
    def z(self, targets, truncate: bool = True): self._one("z", targets, None, truncate)
    def h(self, targets, truncate: bool = True): self._one("h", targets, None, truncate)
    def s(self, targets, truncate: bool = True): self._one("s", targets, None, truncate)
    def s_dag(self, targets, truncate: bool = True): self._one("s_dag", targets, None, truncate)
    def sqrt_x(self, targets, truncate: bool = True): self._one("sqrt_x", targets, None, truncate)
    def sqrt_y(self, targets, truncate: bool = True): self._one("sqrt_y", targets, None, truncate)
    def sqrt_x_dag(self, targets, truncate: bool = True): self._one("sqrt_x_dag", targets, None, truncate)
    def sqrt_y_dag(self, targets, truncate: bool = True): self._one("sqrt_y_dag", targets, None, truncate)

    def cnot(self, targets, truncate: bool = True): self._two("cnot", targets, None, truncate)
    def cz(self, targets, truncate: bool = True): self._two("cz", targets, None, truncate)
    def cy(self, targets, truncate: bool = True): self._two("cy", targets, None, truncate)
    # stim aliases (interface.rs:195-230)
    cx = zcx = cnot
    zcz = cz
    zcy = cy

    def rx(self, targets, theta: float, truncate: bool = True): self._one("rx", targets, float(theta), truncate)
    def ry(self, targets, theta: float, truncate: bool = True): self._one("ry", targets, float(theta), truncate)
    def rz(self, targets, theta: float, truncate: bool = True): self._one("rz", targets, float(theta), truncate)
    def rxx(self, targets, theta: float, truncate: bool = True): self._two("rxx", targets, float(theta), truncate)
    def ryy(self, targets, theta: float, truncate: bool = True): self._two("ryy", targets, float(theta), truncate)
    def rzz(self, targets, theta: float, truncate: bool = True): self._two("rzz", targets, float(theta), truncate)

    def pauli_error(self, targets, p, truncate: bool = True):
        p = [float(v) for v in p]
        if len(p) != 3:
            raise ValueError("pauli_error takes three probabilities (X, Y, Z)")
        for q in targets:
            self._h.pauli_error(int(q), p)
        self._after(truncate)

    def x_error(self, targets, p: float, truncate: bool = True):
        self.pauli_error(targets, [float(p), 0.0, 0.0], truncate)

    def y_error(self, targets, p: float, truncate: bool = True):
        self.pauli_error(targets, [0.0, float(p), 0.0], truncate)

    def z_error(self, targets, p: float, truncate: bool = True):
        self.pauli_error(targets, [0.0, 0.0, float(p)], truncate)

    def two_qubit_pauli_error(self, targets, p, truncate: bool = True):
        p = [float(v) for v in p]
        if len(p) != 15:
            raise ValueError("two_qubit_pauli_error takes fifteen probabilities")
        t = _pairs(targets)
        for a, b in zip(t[0::2], t[1::2]):
            self._h.two_qubit_pauli_error(a, b, p)
        self._after(truncate)

    def depolarize1(self, targets, p: float, truncate: bool = True):
        for q in targets:
            self._h.depolarize1(int(q), float(p))
        self._after(truncate)

    def depolarize2(self, targets, p: float, truncate: bool = True):
        t = _pairs(targets)
        for a, b in zip(t[0::2], t[1::2]):
            self._h.depolarize2(a, b, float(p))
        self._after(truncate)

    def amplitude_damping(self, addr0: int, gamma: float, truncate: bool = True):
        self._h.amplitude_damping(int(addr0), float(gamma))
        self._after(truncate)


class _PauliSumLossBase(_PauliSumBase):
    """``_PauliSumLossBase`` of _core.pyi: lossy words (alphabet I X Y Z L) and
    the loss channels (crates/ppvm-pauli-sum/src/sum/noise.rs:159-274)."""

    _lossy = True

    def _handle(self, n, tau, mw, mlw):
        return _native.LossyHandle(n, tau, mw, None if mlw >= UMAX else mlw)

    def _pack(self, strings):
        return pack_lossy_strings(strings, self.n_qubits)

    def _words(self):
        x, z, l, c = self._h.export()
        return unpack_lossy_words(x, z, l, self.n_qubits), weight_of(np.bitwise_or(x, l), z), c

    def loss_channel(self, addr0: int, p: float, truncate: bool = True):
        self._h.loss_channel(int(addr0), float(p))
        self._after(truncate)

    def correlated_loss_channel(self, addr0: int, addr1: int, p, truncate: bool = True):
        p = [float(v) for v in p]
        if len(p) != 3:
            raise ValueError("correlated_loss_channel takes three probabilities")
        self._h.correlated_loss_channel(int(addr0), int(addr1), p)
        self._after(truncate)

    def reset_loss_channel(self, addr0: int, truncate: bool = True):
        self._h.reset_loss_channel(int(addr0))
        self._after(truncate)


def _tier_class(base, prefix: str, k: int):
    return type(f"{prefix}{k}", (base,), {"_tier_bytes": 1 << k, "__module__": __name__,
                                          "__doc__": f"{base.__doc__.split(':')[0]}: {1 << k} bytes per plane "
                                                     f"(at most {8 << k} qubits)."})


for _k in TIERS:
    globals()[f"PauliSumIndexMapFxHash{_k}"] = _tier_class(_PauliSumBase, "PauliSumIndexMapFxHash", _k)
    globals()[f"PauliSumLossIndexMapFxHash{_k}"] = _tier_class(_PauliSumLossBase, "PauliSumLossIndexMapFxHash", _k)
del _k


def tier_for(n_qubits: int) -> int:
    """The tier the reference wrapper picks: the first k with 2**k above
    ceil(n_qubits / 8) bytes (ppvm-python/src/ppvm/paulisum.py:176-182)."""
    nbytes = -(-int(n_qubits) // 8)
    return next(k for k in TIERS if 2 ** k > nbytes)
