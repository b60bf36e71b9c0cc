"""The learned cost model's network: the FP64 oracle pinned by central
differences and a closed-form network (CPU), then libfot's MFMA GEMM,
forward, gradients and training against the oracle (gpu)."""
import numpy as np
import pytest

import mlp_oracle as MO


def _rand_problem(rng, rows, n_in, hidden, n_out, masked=True):
    p = rng.normal(0, 0.5, hidden * n_in + hidden + hidden * hidden + hidden + n_out * hidden
                   + n_out)
    x = rng.normal(0, 1, (rows, n_in))
    t = rng.normal(0, 1, (rows, n_out))
    m = (rng.random((rows, n_out)) > 0.2).astype(np.float64) if masked else None
    return p, x, t, m


def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    n_in, hidden, n_out = 5, 7, 3
    p, x, t, m = _rand_problem(rng, 20, n_in, hidden, n_out)
    _, g = MO.loss_grad(p, x, t, m, n_in, hidden, n_out)
    idx = rng.choice(p.size, 60, replace=False)
    num = MO.numeric_grad(p, x, t, m, n_in, hidden, n_out, idx)
    np.testing.assert_allclose(g[idx], num, rtol=1e-5, atol=1e-8)


def test_oracle_forward_closed_form():
    # W1 = [I; -I], W2 = I, W3 = [I, -I]: relu(x) - relu(-x) = x exactly
    n = 4
    W1 = np.vstack([np.eye(n), -np.eye(n)])
    W2 = np.eye(2 * n)
    W3 = np.hstack([np.eye(n), -np.eye(n)])
    p = np.concatenate([W1.ravel(), np.zeros(2 * n), W2.ravel(), np.zeros(2 * n), W3.ravel(),
                        np.zeros(n)])
    x = np.random.default_rng(1).normal(size=(9, n))
    y, _ = MO.forward(p, x, n, 2 * n, n)
    np.testing.assert_array_equal(y, x)


# ---- GPU ------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K,ta,tb", [(1, 1, 1, False, False), (17, 5, 3, False, False),
                                         (64, 64, 24, False, True), (100, 6, 64, False, True),
                                         (6, 64, 100, True, False), (33, 70, 19, True, True),
                                         (257, 64, 64, False, True)])
def test_gemm_mfma_matches_numpy(gpu, M, N, K, ta, tb):
    from cuda_fot_amd.mlp import gemm
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    A = a.T.astype(np.float64) if ta else a.astype(np.float64)
    B = b.T.astype(np.float64) if tb else b.astype(np.float64)
    ref = A @ B
    scale = np.abs(A) @ np.abs(B)
    c = gemm(a, b, ta, tb)
    assert c.shape == (M, N)
    assert np.all(np.abs(c - ref) <= 2e-6 * scale + 1e-30), np.abs(c - ref).max()


@pytest.mark.gpu
def test_mlp_forward_and_gradient_match_oracle(gpu):
    from cuda_fot_amd.mlp import MLP
    rng = np.random.default_rng(3)
    net = MLP(24, 64, 6, seed=5)
    p = net.get_params().astype(np.float64)
    x = rng.normal(size=(333, 24)).astype(np.float32)
    t = rng.normal(size=(333, 6)).astype(np.float32)
    m = (rng.random((333, 6)) > 0.3).astype(np.float32)
    y = net.forward(x)
    yr, _ = MO.forward(p, x, 24, 64, 6)
    np.testing.assert_allclose(y, yr, rtol=1e-4, atol=1e-4 * np.abs(yr).max())
    loss, g = net.loss_grad(x, t, m)
    lr, gr = MO.loss_grad(p, x, t, m, 24, 64, 6)
    assert abs(loss - lr) <= 1e-4 * lr
    # per parameter block, relative to the block's largest gradient
    k = 0
    for shape in [(64, 24), (64,), (64, 64), (64,), (6, 64), (6,)]:
        n = int(np.prod(shape))
        gb, rb = g[k:k + n], gr[k:k + n]
        np.testing.assert_allclose(gb, rb, atol=2e-4 * np.abs(rb).max() + 1e-12, rtol=0)
        k += n


@pytest.mark.gpu
def test_mlp_training_fits_a_smooth_target(gpu):
    from cuda_fot_amd.mlp import MLP
    rng = np.random.default_rng(4)
    x = rng.normal(size=(512, 24)).astype(np.float32)
    t = np.stack([np.sin(x[:, i]) + 0.5 * x[:, i + 6] for i in range(6)], 1).astype(np.float32)
    net = MLP(24, 64, 6, seed=1)
    losses = net.train(x, t, epochs=600, lr=3e-3)
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
    # the trained parameters reproduce the reported loss through the oracle
    lr_, _ = MO.loss_grad(net.get_params().astype(np.float64), x, t, None, 24, 64, 6)
    assert lr_ < 0.1 * losses[0]
