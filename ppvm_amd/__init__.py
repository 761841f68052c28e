"""gfx950 (MI355X) Pauli-propagation engine with the PauliSum surface of
QuEraComputing/ppvm.  See include/ppvm_hip.h for the C ABI."""

from .ops import Program  # noqa: F401
from .paulisum import LossyPauliSum, PauliSum  # noqa: F401

__all__ = ["LossyPauliSum", "PauliSum", "Program"]
