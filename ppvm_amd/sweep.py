"""Parameter sweeps of the 10x10 lattice TFIM with the GPU propagator.

Prefix reuse: applying the same backward Trotter step m times gives the
Heisenberg operator O_m, so recording overlap_with_zero after every step of a
single propagation yields <Z_(5,5)>(m dt) for all m at once.  A sweep over
(h, t) therefore needs one propagation per h value, for the largest step
count in the grid.  Grid of the surrogate workload: 64 h values in
[0.5, 4.0] x 64 t values in [0, 2.0], dt = 0.05, steps = round(t / dt).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import workloads as WL

DT = 0.05
H_GRID = np.linspace(0.5, 4.0, 64)
T_GRID = np.linspace(0.0, 2.0, 64)


def steps_for(t: np.ndarray, dt: float = DT) -> np.ndarray:
    return np.rint(np.asarray(t) / dt).astype(np.int64)


def lattice_series(h: float, steps: int, capacity_hint: int = 0):
    """Overlap after 0..steps backward Trotter steps (index m = after m steps).

    Returns (overlap[m], live terms[m], term updates, seconds[m]) where
    seconds[m] is the wall time from the start of the call to the overlap
    after m steps (each overlap is read back, so the GPU work is done)."""
    from .paulisum import PauliSum
    t0 = time.perf_counter()
    w = WL.lattice_tfim(h=h, steps=steps)
    g = PauliSum.new(w.n_qubits, w.observable, min_abs_coeff=w.min_abs_coeff,
                     max_pauli_weight=w.max_pauli_weight, capacity_hint=capacity_hint)
    out = [g.overlap_with_zero()]
    lens = [len(g)]
    secs = [time.perf_counter() - t0]
    prev = 0
    for li, m in enumerate(w.layer_marks):
        g.apply(w.program[prev:m + 1])
        prev = m + 1
        if li % 3 == 2:  # three layers (rx, vertical rzz, horizontal rzz) per step
            out.append(g.overlap_with_zero())
            lens.append(len(g))
            secs.append(time.perf_counter() - t0)
    tu = g.stats()["term_updates"]
    return np.array(out), np.array(lens), tu, np.array(secs)


def plain_sweep(h_values, steps, log=None):
    """One propagation per h value, to ``steps`` (an int, or one count per h).
    Returns (series per h, cumulative seconds per h, term updates per h)."""
    steps = np.broadcast_to(np.asarray(steps, dtype=np.int64), (len(h_values),))
    series, cum = [], []
    tus = np.zeros(len(h_values), dtype=np.int64)
    for i, h in enumerate(h_values):
        ser, lens, tu, secs = lattice_series(float(h), int(steps[i]))
        series.append(ser)
        cum.append(secs)
        tus[i] = tu
        if log:
            log(f"h={h:.4f}: {int(steps[i])} steps {secs[-1]:.2f}s peak {int(lens.max())} terms")
    return series, cum, tus


def main() -> int:
    ap = argparse.ArgumentParser(description="lattice TFIM sweep with the GPU propagator")
    ap.add_argument("--out", required=True)
    ap.add_argument("--steps", type=int, default=int(steps_for(T_GRID).max()))
    args = ap.parse_args()
    series, cum, tus = plain_sweep(H_GRID, args.steps,
                                   log=lambda m: print(m, file=sys.stderr, flush=True))
    grid = np.stack(series)
    secs = np.array([c[-1] for c in cum])
    np.savez(args.out, h=H_GRID, t=T_GRID, dt=DT, grid=grid, seconds=secs, term_updates=tus)
    print(json.dumps({"h_values": len(H_GRID), "steps": args.steps, "total_s": float(secs.sum()),
                      "term_updates": int(tus.sum())}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
