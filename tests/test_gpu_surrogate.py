"""What the learned surrogate gains over the plain propagator sweep.

The ensemble is trained on the GPU on 2048 seed-31 points of the shipped
propagator grid (10x10 lattice TFIM, 64 h x 64 t) and queried on the other
2048.  Cost of the plain method for the query set: one propagation per h value
that holds a query point (prefix reuse), at the wall times recorded with the
grid.  Cost of the surrogate: inference plus one propagation per h value where
the ensemble std exceeds 1e-2.  The test checks the gain (>= 10x) and the
error distribution against the propagator values.
"""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_gain_over_plain_sweep():
    from ppvm_amd.surrogate import load_grid, run_config5
    d = load_grid()
    t0 = time.perf_counter()
    rep = run_config5(d["grid"], d["h"], d["t"], d["seconds"], steps=12000, lr=2e-3)
    print(f"train {rep.train_s:.2f}s infer {rep.infer_s * 1e3:.2f} ms fallback h {rep.fallback_h} "
          f"frac<=1e-3 {rep.frac_within_1e3:.3f} median {np.median(rep.abs_err):.2e} "
          f"max {rep.max_err:.2e} speedup {rep.speedup:.1f} ({time.perf_counter() - t0:.1f}s)")
    assert rep.n_query == 2048
    assert rep.speedup >= 10.0
    assert np.median(rep.abs_err) < 2e-3
    assert rep.frac_within_1e3 > 0.4
    assert rep.max_err < 0.1
