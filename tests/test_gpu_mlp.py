"""Surrogate MLP kernels (BF16 MFMA, FP32 accumulation) vs the numpy reference.

* the MFMA GEMM (lane layout, tiling, edges) against a float64 product of the
  same BF16-rounded operands;
* the gradients of one forward + backward pass against oracle/mlp_ref.py
  (same rounding points): relative L2 error per layer below 2e-2;
* training lowers the loss and the ensemble statistics match the members.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (70, 40, 48), (257, 130, 300), (2048, 32, 256)])
def test_mfma_gemm_matches_numpy(M, N, K):
    from oracle.mlp_ref import bf16
    from ppvm_amd.surrogate import gemm_check
    rng = np.random.default_rng(M * 7 + N)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(N, K)).astype(np.float32)
    C = gemm_check(A, B)
    ref = bf16(A).astype(np.float64) @ bf16(B).astype(np.float64).T
    err = np.max(np.abs(C - ref)) / np.max(np.abs(ref))
    assert err < 1e-5, err


@pytest.mark.parametrize("hidden", [64, 256])
def test_gradients_match_reference(hidden):
    from oracle.mlp_ref import forward_backward, unpack
    from ppvm_amd.surrogate import MLPEnsemble
    rng = np.random.default_rng(hidden)
    n, d = 300, 11
    X = rng.uniform(-1, 1, size=(n, d)).astype(np.float32)
    Y = np.sin(2 * X[:, 0]) * np.cos(X[:, 1])
    net = MLPEnsemble(d, hidden, 2, 512, seed=5)
    for member in range(2):
        p = net.params(member)
        loss_g, g = net.grads(X, Y, member)
        loss_r, gr = forward_backward(p, X, Y, hidden)
        assert abs(loss_g - loss_r) <= 1e-3 * max(1.0, loss_r)
        for (Wg, bg), (Wr, br) in zip(unpack(g, hidden), unpack(gr, hidden)):
            for a, b in ((Wg, Wr), (bg, br)):
                nb = np.linalg.norm(b)
                if nb == 0:
                    assert np.linalg.norm(a) == 0
                    continue
                assert np.linalg.norm(a - b) / nb < 2e-2


def test_training_reduces_loss_and_ensemble_stats():
    from ppvm_amd.surrogate import MLPEnsemble
    rng = np.random.default_rng(2)
    n, d = 1024, 6
    X = rng.uniform(-1, 1, size=(n, d)).astype(np.float32)
    Y = (0.5 * np.sin(3 * X[:, 0]) + 0.3 * X[:, 1] * X[:, 2]).astype(np.float32)
    net = MLPEnsemble(d, 128, 4, 2048, seed=9)
    l0 = net.train(X, Y, steps=1, lr=1e-3)
    l1 = net.train(X, Y, steps=600, lr=2e-3)
    assert np.all(l1 < 0.1 * l0), (l0, l1)
    mean, sd, pm = net.predict(X, per_member=True)
    assert np.allclose(mean, pm.mean(axis=0), atol=1e-6)
    assert np.allclose(sd, pm.std(axis=0), atol=1e-5)
    assert np.sqrt(np.mean((mean - Y) ** 2)) < 0.05
