"""``PauliSum`` with the method surface of the reference Python package
(ppvm-python/src/ppvm/paulisum.py:80-416 and the Truncating mixins of
ppvm-python/src/ppvm/mixins.py:523-870), backed by the gfx950 engine.

Same constructor, defaults and cadence as the reference: ``min_abs_coeff``
defaults to 1e-10, ``max_pauli_weight=None`` disables the weight cap, and every
gate / channel call truncates afterwards unless ``truncate=False`` is passed
(crates/ppvm-python-native/src/interface.rs:151-360).  Gates must be applied in
reverse circuit order (Heisenberg picture).
"""

from __future__ import annotations

from collections.abc import Iterable, Sequence
from typing import Any

import numpy as np

from . import _native
from .ops import pack_lossy_strings, pack_strings, unpack_lossy_words, unpack_words, weight_of

_TWO_Q_ORDER = ("IX", "IY", "IZ", "XI", "XX", "XY", "XZ", "YI", "YX", "YY", "YZ", "ZI", "ZX",
                "ZY", "ZZ")


def _compact_tokens(s: str):
    """Split the compact form ``"X3Z5"`` (a Pauli letter and a qubit index,
    repeated, nothing else) into ``[("X", 3), ("Z", 5)]``; None if ``s`` is not
    in that form (then it is a full string)."""
    tokens, i, n = [], 0, len(s)
    if n == 0:
        return None
    while i < n:
        letter = s[i]
        if letter not in "IXYZ":
            return None
        j = i + 1
        while j < n and s[j].isdecimal():
            j += 1
        if j == i + 1:
            return None
        tokens.append((letter, int(s[i + 1:j])))
        i = j
    return tokens


def _parse_term(term, n_qubits: int):
    """A term given as a full string, in the compact form, or as a
    ``(string, coefficient)`` tuple, as the reference's PauliSum.new accepts
    (ppvm-python/src/ppvm/paulisum.py:20-38): returns (full string, coefficient).
    In the compact form, unnamed qubits are I and a later letter for the same
    qubit wins."""
    text, coeff = term if isinstance(term, tuple) else (term, 1.0)
    tokens = _compact_tokens(text)
    if tokens is not None:
        full = ["I"] * n_qubits
        for letter, q in tokens:
            if q >= n_qubits:
                raise ValueError(f"Qubit index {q} out of range for {n_qubits}-qubit system.")
            full[q] = letter
        text = "".join(full)
    return text, float(coeff)


def _is_sequence(obj: Any) -> bool:
    if isinstance(obj, (list, tuple)):
        return True
    if isinstance(obj, (int, str, bytes)):
        return False
    return isinstance(obj, Iterable)


def _targets(args) -> list[int]:
    if len(args) == 1 and _is_sequence(args[0]):
        return [int(t) for t in args[0]]
    return [int(t) for t in args]


def _split(args, value, name, truncate):
    """(*targets, value[, truncate]) splitting of mixins.py:96-124."""
    if args and _is_sequence(args[0]):
        targets = list(args[0])
        rest = list(args[1:])
        if rest and isinstance(rest[-1], bool):
            truncate = rest.pop()
        if value is None:
            if not rest:
                raise TypeError(f"missing required argument: {name!r}")
            value = rest[-1]
        return [int(t) for t in targets], value, truncate
    if value is None:
        if not args:
            raise TypeError(f"missing required argument: {name!r}")
        a = list(args)
        if len(a) >= 2 and isinstance(a[-1], bool):
            truncate = a.pop()
        return [int(t) for t in a[:-1]], a[-1], truncate
    return [int(t) for t in args], value, truncate


def _rust_fixed3(v: float) -> str:
    """Rust's `{:.3}` for f64."""
    if v != v:
        return "NaN"
    if v in (float("inf"), float("-inf")):
        return "inf" if v > 0 else "-inf"
    return f"{v:.3f}"


class PauliSum:
    """A weighted sum of Pauli strings propagated on the GPU.

    Args mirror the reference dataclass: ``initial_terms`` (strings of length
    ``n_qubits``), ``coefficients`` (default all 1.0), ``min_abs_coeff``,
    ``max_pauli_weight``, ``preserve_strings``.
    """

    def __init__(self, initial_terms: Sequence[str], n_qubits: int,
                 coefficients: Sequence[float] = (), min_abs_coeff: float = 1e-10,
                 max_pauli_weight: int | None = None, max_loss_weight: int | None = None,
                 preserve_strings: Sequence[str] | None = None, capacity_hint: int = 0):
        if not initial_terms:
            raise ValueError("At least one term must be provided to initialize PauliSum.")
        for term in initial_terms:
            if len(term) != n_qubits:
                raise ValueError(
                    "All terms must have the same length! Expected length "
                    f"{n_qubits}, but got term of length {len(term)}: {term!r}")
        coefficients = tuple(float(c) for c in coefficients)
        if not coefficients:
            coefficients = (1.0,) * len(initial_terms)
        if len(coefficients) != len(initial_terms):
            raise ValueError("Initial terms and coefficients need to be of same length!")
        self.initial_terms = tuple(initial_terms)
        self.n_qubits = int(n_qubits)
        self.coefficients = coefficients
        self.min_abs_coeff = float(min_abs_coeff)
        self.max_pauli_weight = max_pauli_weight
        self.max_loss_weight = max_loss_weight
        self.preserve_strings = tuple(preserve_strings) if preserve_strings else None
        self._h = self._new_handle(capacity_hint)
        if self.preserve_strings:
            for s in self.preserve_strings:
                if len(s) != n_qubits:
                    raise ValueError("All preserve strings must have length n_qubits "
                                     f"({n_qubits}); got {len(s)}: {s!r}")
            self._h.set_preserve(*self._pack(self.preserve_strings))
        self._h.add_words(*self._pack(self.initial_terms), np.asarray(coefficients, dtype=np.float64))

    def _new_handle(self, capacity_hint: int):
        # max_loss_weight only applies to lossy words; like the reference's
        # non-lossy classes, this one accepts and ignores it (interface.rs:89)
        return _native.Handle(self.n_qubits, self.min_abs_coeff, self.max_pauli_weight, capacity_hint)

    def _pack(self, strings):
        return pack_strings(strings, self.n_qubits)

    @classmethod
    def new(cls, n_qubits: int, terms, min_abs_coeff: float = 1e-10,
            max_pauli_weight: int | None = None, max_loss_weight: int | None = None,
            preserve_strings: Sequence[str] | None = None, capacity_hint: int = 0) -> "PauliSum":
        if isinstance(terms, (str, tuple)):
            terms = [terms]
        parsed = [_parse_term(t, n_qubits) for t in terms]
        return cls(initial_terms=[s for s, _ in parsed], n_qubits=n_qubits,
                   coefficients=[c for _, c in parsed], min_abs_coeff=min_abs_coeff,
                   max_pauli_weight=max_pauli_weight, max_loss_weight=max_loss_weight,
                   preserve_strings=preserve_strings, capacity_hint=capacity_hint)

    def reset(self) -> None:
        """Restore the initial terms on the same device arenas (no reallocation):
        the sum is cleared and the constructor's terms are added again."""
        self._h.clear()
        self._h.add_words(*self._pack(self.initial_terms), np.asarray(self.coefficients, dtype=np.float64))

    # ---- bookkeeping --------------------------------------------------------------
    def __len__(self) -> int:
        return len(self._h)

    def __copy__(self) -> "PauliSum":
        new = object.__new__(type(self))
        new.__dict__.update({k: v for k, v in self.__dict__.items() if k != "_h"})
        new._h = self._h.clone()
        return new

    def copy(self) -> "PauliSum":
        return self.__copy__()

    def __deepcopy__(self, memo) -> "PauliSum":
        return self.__copy__()

    def _export(self):
        return self._h.export()

    @property
    def terms(self) -> list[tuple[str, float]]:
        x, z, c = self._export()
        return list(zip(unpack_words(x, z, self.n_qubits), c.tolist()))

    def weights(self) -> list[tuple[str, int]]:
        x, z, _ = self._export()
        return list(zip(unpack_words(x, z, self.n_qubits), weight_of(x, z).tolist()))

    def current_max_weight(self) -> int:
        x, z, _ = self._export()
        return int(weight_of(x, z).max()) if len(x) else 0

    def __eq__(self, other) -> bool:
        if not isinstance(other, PauliSum) or other.n_qubits != self.n_qubits:
            return NotImplemented
        a = dict(self.terms)
        b = dict(other.terms)
        return a == b

    def __str__(self) -> str:
        # "{:.3} * {word}" joined by " + ", stably sorted by weight
        # (crates/ppvm-pauli-sum/src/sum/display.rs:29-44); Rust's {:.3} is fixed-point
        w = self.weights()
        t = dict(self.terms)
        rows = sorted(w, key=lambda sw: sw[1])
        return " + ".join(f"{_rust_fixed3(t[s])} * {s}" for s, _ in rows)

    # ---- observables --------------------------------------------------------------
    def overlap_with_zero(self) -> float:
        """<0...0| O |0...0>: sum of coefficients of I/Z-only terms."""
        return self._h.overlap_with_zero()

    def trace(self, pattern: str) -> float:
        """Sum of the coefficients of terms matching a Pauli pattern
        (crates/ppvm-python-native/src/interface.rs:122-126): literal X/Y/Z,
        '_', [..] sets, '?' (identity allowed), '*', '{n}' and qubit anchors.
        ValueError if the pattern does not parse."""
        if pattern.strip() == "Z?*":
            return self.overlap_with_zero()
        return self._h.trace(pattern)

    def overlap(self, other: "PauliSum") -> float:
        """sum_k other[k] * self[k] (crates/ppvm-pauli-sum/src/sum/trace.rs:30-37)."""
        if type(other._h) is not type(self._h):
            raise TypeError("overlap needs two sums of the same kind")
        return self._h.overlap(other._h)

    def truncate(self) -> None:
        self._h.truncate()

    def synchronize(self) -> None:
        self._h.synchronize()

    def stats(self) -> dict:
        return self._h.stats()

    # ---- gate helpers ----------------------------------------------------------------
    def _g(self, name, targets, theta=None, truncate=True):
        self._h.gate(name, targets, theta)
        if truncate:
            self._h.truncate()

    # Cliffords (TruncatingCliffordMixin / TruncatingCliffordExtensionMixin)
    def x(self, *targets, truncate: bool = True): self._g("x", _targets(targets), None, truncate)
    def y(self, *targets, truncate: bool = True): self._g("y", _targets(targets), None, truncate)
    def z(self, *targets, truncate: bool = True): self._g("z", _targets(targets), None, truncate)
    def h(self, *targets, truncate: bool = True): self._g("h", _targets(targets), None, truncate)
    def s(self, *targets, truncate: bool = True): self._g("s", _targets(targets), None, truncate)
    def s_dag(self, *targets, truncate: bool = True): self._g("s_dag", _targets(targets), None, truncate)
    def sqrt_x(self, *targets, truncate: bool = True): self._g("sqrt_x", _targets(targets), None, truncate)
    def sqrt_x_dag(self, *targets, truncate: bool = True): self._g("sqrt_x_dag", _targets(targets), None, truncate)
    def sqrt_y(self, *targets, truncate: bool = True): self._g("sqrt_y", _targets(targets), None, truncate)
    def sqrt_y_dag(self, *targets, truncate: bool = True): self._g("sqrt_y_dag", _targets(targets), None, truncate)
    def cnot(self, *targets, truncate: bool = True): self._g("cnot", _targets(targets), None, truncate)
    def cz(self, *targets, truncate: bool = True): self._g("cz", _targets(targets), None, truncate)
    def cy(self, *targets, truncate: bool = True): self._g("cy", _targets(targets), None, truncate)
    cx = cnot
    zcx = cnot
    zcz = cz

    # Rotations (TruncatingRotationsMixin)
    def rx(self, *args, theta=None, truncate: bool = True):
        t, th, tr = _split(args, theta, "theta", truncate)
        self._g("rx", t, th, tr)

    def ry(self, *args, theta=None, truncate: bool = True):
        t, th, tr = _split(args, theta, "theta", truncate)
        self._g("ry", t, th, tr)

    def rz(self, *args, theta=None, truncate: bool = True):
        t, th, tr = _split(args, theta, "theta", truncate)
        self._g("rz", t, th, tr)

    def rxx(self, *args, theta=None, truncate: bool = True):
        t, th, tr = _split(args, theta, "theta", truncate)
        self._g("rxx", t, th, tr)

    def ryy(self, *args, theta=None, truncate: bool = True):
        t, th, tr = _split(args, theta, "theta", truncate)
        self._g("ryy", t, th, tr)

    def rzz(self, *args, theta=None, truncate: bool = True):
        t, th, tr = _split(args, theta, "theta", truncate)
        self._g("rzz", t, th, tr)

    def r(self, addr0: int, axis_angle: float, theta: float, *, truncate: bool = True):
        """R(phi, theta) = rz(phi); rx(theta); rz(-phi) (crates/ppvm-pauli-sum/src/sum/rot1.rs:135-143)."""
        self._h.r(int(addr0), float(axis_angle), float(theta))
        if truncate:
            self._h.truncate()

    # Noise (TruncatingNoiseMixin)
    def pauli_error(self, *args, p=None, truncate: bool = True):
        t, pv, tr = _split(args, p, "p", truncate)
        for q in t:
            self._h.pauli_error(int(q), pv)
        if tr:
            self._h.truncate()

    def x_error(self, *args, p=None, truncate: bool = True):
        t, pv, tr = _split(args, p, "p", truncate)
        self.pauli_error(t, [pv, 0.0, 0.0], truncate=tr)

    def y_error(self, *args, p=None, truncate: bool = True):
        t, pv, tr = _split(args, p, "p", truncate)
        self.pauli_error(t, [0.0, pv, 0.0], truncate=tr)

    def z_error(self, *args, p=None, truncate: bool = True):
        t, pv, tr = _split(args, p, "p", truncate)
        self.pauli_error(t, [0.0, 0.0, pv], truncate=tr)

    def depolarize1(self, *args, p=None, truncate: bool = True):
        t, pv, tr = _split(args, p, "p", truncate)
        for q in t:
            self._h.depolarize1(int(q), float(pv))
        if tr:
            self._h.truncate()

    def amplitude_damping(self, addr0: int, gamma: float, *, truncate: bool = True):
        """Amplitude damping (ppvm-python/src/ppvm/paulisum.py:389-403): X and Y on
        the qubit are damped by sqrt(1 - gamma); Z branches into
        gamma * I + (1 - gamma) * Z."""
        self._h.amplitude_damping(int(addr0), float(gamma))
        if truncate:
            self._h.truncate()

    def depolarize2(self, *args, p=None, truncate: bool = True):
        t, pv, tr = _split(args, p, "p", truncate)
        if len(t) % 2:
            raise ValueError("Expected an even number of targets for a two-qubit channel")
        for a, b in zip(t[0::2], t[1::2]):
            self._h.depolarize2(int(a), int(b), float(pv))
        if tr:
            self._h.truncate()

    @staticmethod
    def two_qubit_pauli_error_probabilities(error_probabilities: dict) -> list[float]:
        return [float(error_probabilities.get(k, 0.0)) for k in _TWO_Q_ORDER]

    def two_qubit_pauli_error(self, *args, p=None, truncate: bool = True):
        t, pv, tr = _split(args, p, "p", truncate)
        if len(t) % 2:
            raise ValueError("Expected an even number of targets for a two-qubit channel")
        for a, b in zip(t[0::2], t[1::2]):
            self._h.two_qubit_pauli_error(int(a), int(b), pv)
        if tr:
            self._h.truncate()

    # ---- programs ---------------------------------------------------------------------
    def apply(self, program) -> None:
        """Apply a whole op program (``ppvm_amd.ops.Program`` or an OP_DTYPE array)."""
        arr = program.array() if hasattr(program, "array") else program
        self._h.apply(arr)


class LossyPauliSum(PauliSum):
    """A PauliSum over lossy words: the Pauli alphabet {I, X, Y, Z, L}, where
    ``L`` marks a lost qubit (ppvm-python/src/ppvm/paulisum.py:420-520,
    crates/ppvm-pauli-word/src/loss/data.rs).  Gates skip lost qubits as the
    reference does; ``max_loss_weight`` drops terms with more lost qubits than
    that; the loss channels below are available on this class only.
    """

    def _new_handle(self, capacity_hint: int):
        return _native.LossyHandle(self.n_qubits, self.min_abs_coeff, self.max_pauli_weight,
                                   self.max_loss_weight, capacity_hint)

    def _pack(self, strings):
        return pack_lossy_strings(strings, self.n_qubits)

    @property
    def terms(self) -> list[tuple[str, float]]:
        x, z, l, c = self._export()
        return list(zip(unpack_lossy_words(x, z, l, self.n_qubits), c.tolist()))

    def weights(self) -> list[tuple[str, int]]:
        x, z, l, _ = self._export()
        return list(zip(unpack_lossy_words(x, z, l, self.n_qubits),
                        weight_of(np.bitwise_or(x, l), z).tolist()))

    def current_max_weight(self) -> int:
        x, z, l, _ = self._export()
        return int(weight_of(np.bitwise_or(x, l), z).max()) if len(x) else 0

    def loss_weights(self) -> list[tuple[str, int]]:
        x, z, l, _ = self._export()
        return list(zip(unpack_lossy_words(x, z, l, self.n_qubits), weight_of(l, l).tolist()))

    def trace(self, pattern: str) -> float:
        return self._h.trace(pattern)

    def loss_channel(self, addr0: int, p: float, *, truncate: bool = True) -> None:
        """Single-qubit loss (crates/ppvm-pauli-sum/src/sum/noise.rs:159-180): a
        term with L on the qubit adds p * (word with I there); any other term is
        scaled by (1 - p)."""
        self._h.loss_channel(int(addr0), float(p))
        if truncate:
            self._h.truncate()

    def correlated_loss_channel(self, addr0: int, addr1: int, p: Sequence[float], *,
                                truncate: bool = True) -> None:
        """Correlated loss of two qubits (noise.rs:182-248); ``p`` = (both lost,
        one of two lost, the second lost once the first is)."""
        if len(p) != 3:
            raise ValueError("correlated_loss_channel takes three probabilities")
        self._h.correlated_loss_channel(int(addr0), int(addr1), [float(v) for v in p])
        if truncate:
            self._h.truncate()

    def reset_loss_channel(self, addr0: int, *, truncate: bool = True) -> None:
        """Reset a lost qubit to |0> (noise.rs:250-274): I and Z terms add the
        word with L on the qubit; L terms are zeroed."""
        self._h.reset_loss_channel(int(addr0))
        if truncate:
            self._h.truncate()
