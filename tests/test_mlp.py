"""Planner MLP: the float64 CPU oracle is pinned by central finite
differences (CPU); the GPU forward/backward kernels are checked against it
(GPU); training on exact-pass tile statistics lowers the loss (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _data(rng, n=96):
    x = rng.normal(size=(n, M.NI)) * 2 + 3
    t = np.stack([np.sin(x[:, 0]) + x[:, 3] * 0.5, np.cos(x[:, 1]) * 2 + x[:, 2] * 0.1], 1)
    return x, t


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    x, t = _data(rng, 40)
    p = M.random_params(rng, x, t)
    loss, g = M.loss_and_grad(p, x, t)
    assert loss == pytest.approx(M.loss_only(p, x, t), rel=1e-12)
    idx = rng.choice(M.N_TRAIN, size=200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (M.loss_only(p + e, x, t) - M.loss_only(p - e, x, t)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=1e-8), f"param {i}"
    # normalisation constants are not trained
    assert np.all(g[M.N_TRAIN:] == 0)


def test_oracle_detects_planted_mistake():
    """A plausible backprop mistake (ReLU mask dropped in layer 1) fails the
    finite-difference check, so the check has teeth."""
    rng = np.random.default_rng(1)
    x, t = _data(rng, 30)
    p = M.random_params(rng, x, t)
    P = M.unpack(p)
    y, (xn, z1, h1, z2, h2) = M.forward(p, x)
    tn = (t - P["out_mean"]) / P["out_std"]
    dy = 2.0 * (y - tn) / (y.shape[0] * M.NO)
    dz2 = (dy @ P["W3"].T) * (z2 > 0)
    wrong_W1 = xn.T @ (dz2 @ P["W2"].T)  # mask omitted
    right_W1 = M.loss_and_grad(p, x, t)[1][:M.NI * M.NH].reshape(M.NI, M.NH)
    assert not np.allclose(wrong_W1, right_W1, rtol=1e-3, atol=1e-8)


@pytest.mark.gpu
def test_gpu_forward_backward_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(2)
    for n in (1, 63, 64, 200):
        x, t = _data(rng, n)
        p = M.random_params(rng, x, t).astype(np.float32)
        gpu_ctx.set_mlp_params(p)
        pred, loss, grad = gpu_ctx.mlp_forward_backward(x.astype(np.float32), t.astype(np.float32))
        p64 = p.astype(np.float64)
        y, _ = M.forward(p64, x.astype(np.float32).astype(np.float64))
        rl, rg = M.loss_and_grad(p64, x.astype(np.float32).astype(np.float64), t.astype(np.float32).astype(np.float64))
        # FP32 kernels against the float64 oracle: tolerance set by float32 rounding
        np.testing.assert_allclose(pred, y, rtol=2e-4, atol=2e-4)
        assert loss == pytest.approx(rl, rel=2e-4, abs=1e-6)
        scale = np.max(np.abs(rg[:M.N_TRAIN])) + 1e-12
        np.testing.assert_allclose(grad[:M.N_TRAIN], rg[:M.N_TRAIN], rtol=2e-3, atol=2e-4 * scale)


@pytest.mark.gpu
def test_gpu_training_reduces_loss(gpu_ctx):
    rng = np.random.default_rng(3)
    x, t = _data(rng, 2000)
    loss = gpu_ctx.train_mlp(x.astype(np.float32), t.astype(np.float32), epochs=40, batch=128, lr=3e-3, seed=5)
    p = gpu_ctx.mlp_params().astype(np.float64)
    ref = M.loss_only(p, x.astype(np.float32), t.astype(np.float32))
    assert loss == pytest.approx(ref, rel=1e-3, abs=1e-4)
    assert loss < 0.1  # targets are smooth functions of the features; untrained loss is ~1
