"""The device network (BF16 matrix-core GEMMs, FP32 accumulation) against the
CPU oracle (oracle/mlp.py, which rounds to BF16 where the device does and
accumulates in FP64).

A one-layer network is a bare GEMM plus bias: its output must agree with the
FP64 product of the same BF16 operands to the FP32 accumulation bound, which
fixes the MFMA operand and result layouts, the tile edges and the K loop.
Its gradients fix the two transposed GEMM forms; a two-layer network adds the
GELU derivative; the full 4-256-256-256-64 network is checked end to end.
"""
import numpy as np
import pytest

from nimblesm_amd.surrogate import MLP
from oracle import mlp as M

pytestmark = pytest.mark.gpu

U = 2.0 ** -24  # FP32 unit roundoff


def data(widths, batch, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (batch, widths[0])).astype(np.float32)
    t = rng.normal(0, 1, (batch, widths[-1])).astype(np.float32)
    return x, t


@pytest.mark.parametrize("K,N,B", [(4, 256, 512), (32, 64, 1), (37, 70, 100), (256, 256, 512), (300, 64, 129),
                                   (256, 64, 8193)])
def test_single_layer_is_an_exact_bf16_gemm(hip_device, K, N, B):
    net = MLP((K, N), seed=K + N)
    rng = np.random.default_rng(B)
    p = net.get_parameters()
    p[-N:] = rng.normal(0, 1, N).astype(np.float32)  # nonzero bias
    net.set_parameters(p)
    x, _ = data((K, N), B, K)
    y = net.forward(x).astype(np.float64)
    W = M.to_bf16(p[:N * K].reshape(N, K))
    X = M.to_bf16(x)
    ref = X @ W.T + p[N * K:].astype(np.float64)
    bound = 2 * K * U * (np.abs(X) @ np.abs(W).T) + 4 * U * np.abs(ref)
    assert np.all(np.abs(y - ref) <= bound), np.max(np.abs(y - ref) / bound)
    net.close()


def test_batch_past_65535_row_tiles(hip_device):
    """2,097,192 rows: more 32-row tiles than a grid's y or z dimension may
    hold; rows from every part of the batch, the last ones included, are
    the exact BF16 product."""
    K, N, B = 4, 16, 65536 * 32 + 40
    net = MLP((K, N), seed=5)
    rng = np.random.default_rng(8)
    p = net.get_parameters()
    p[-N:] = rng.normal(0, 1, N).astype(np.float32)
    net.set_parameters(p)
    x = rng.normal(0, 1, (B, K)).astype(np.float32)
    y = net.forward(x).astype(np.float64)
    rows = np.concatenate([np.arange(64), rng.choice(B, 2000, replace=False), np.arange(B - 100, B)])
    W = M.to_bf16(p[:N * K].reshape(N, K))
    X = M.to_bf16(x[rows])
    ref = X @ W.T + p[N * K:].astype(np.float64)
    bound = 2 * K * U * (np.abs(X) @ np.abs(W).T) + 4 * U * np.abs(ref)
    assert np.all(np.abs(y[rows] - ref) <= bound), np.max(np.abs(y[rows] - ref) / bound)
    net.close()


def test_initial_parameters_match_the_oracle(hip_device):
    widths = (4, 256, 256, 256, 64)
    net = MLP(widths, seed=12345)
    assert np.array_equal(net.get_parameters().astype(np.float64), M.init_params(widths, 12345))
    net.close()


@pytest.mark.parametrize("widths,B", [((37, 70), 100), ((4, 256), 512), ((4, 48, 40), 96), ((16, 96, 80), 200),
                                      ((37, 70), 2500)])  # weight gradients over 3 K ranges
def test_gradients_small_networks(hip_device, widths, B):
    net = MLP(widths, seed=7)
    x, t = data(widths, B, 1)
    loss, g = net.loss_gradients(x, t)
    p = net.get_parameters().astype(np.float64)
    rl, rg = M.loss_and_grad(p, widths, x, t, bf16=True)
    assert abs(loss - rl) <= 1e-5 * rl
    # one-layer: D^T H and column sums of G, both of exact BF16 operands
    tol = 2e-5 if len(widths) == 2 else 2e-3
    assert np.linalg.norm(g - rg) / np.linalg.norm(rg) <= tol
    net.close()


def test_full_network_forward_and_gradients(hip_device):
    widths = (4, 256, 256, 256, 64)
    net = MLP(widths, seed=3)
    x, t = data(widths, 512, 2)
    p = net.get_parameters().astype(np.float64)
    y = net.forward(x).astype(np.float64)
    y16, _ = M.forward(p, widths, x, bf16=True)
    y64, _ = M.forward(p, widths, x)
    # BF16 rounding points flip on a few elements between FP32 and FP64
    # accumulation; against the FP64 network the BF16 error itself shows
    assert np.linalg.norm(y - y16) / np.linalg.norm(y16) <= 2e-3
    assert np.linalg.norm(y - y64) / np.linalg.norm(y64) <= 3e-2
    loss, g = net.loss_gradients(x, t)
    rl, rg = M.loss_and_grad(p, widths, x, t, bf16=True)
    assert abs(loss - rl) <= 1e-3 * rl
    assert np.linalg.norm(g - rg) / np.linalg.norm(rg) <= 5e-3
    _, g64 = M.loss_and_grad(p, widths, x, t)
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) <= 5e-2
    net.close()


def test_adam_step(hip_device):
    """One training step is Adam (beta 0.9, 0.999, eps 1e-8) on the gradient."""
    widths = (4, 48, 16)
    net = MLP(widths, seed=9)
    x, t = data(widths, 64, 4)
    p0 = net.get_parameters().astype(np.float64)
    _, g = net.loss_gradients(x, t)
    g = g.astype(np.float64)
    lr = 1e-2
    net.train(x, t, 1, lr, lr)
    p1 = net.get_parameters().astype(np.float64)
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p0 - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.max(np.abs(p1 - ref)) <= 1e-5 * lr + 1e-6 * np.max(np.abs(p0))
    net.close()


def test_training_fits_a_smooth_map(hip_device):
    widths = (4, 256, 256, 256, 64)
    rng = np.random.default_rng(5)
    x = rng.normal(0, 1, (512, 4)).astype(np.float32)
    A = rng.normal(0, 0.5, (4, 64))
    t = (np.sin(x @ A) + 0.1 * (x @ A) ** 2).astype(np.float32)
    net = MLP(widths, seed=1)
    hist = net.train(x, t, 1500, 2e-3, 2e-5)
    assert hist[-1] < 1e-2 * hist[0], (hist[0], hist[-1])
    y = net.forward(x)
    assert np.linalg.norm(y - t) / np.linalg.norm(t) < 0.05
    net.close()


def test_loss_history_across_calls_and_chunks(hip_device):
    """Losses are formed from per-iteration partials in chunks of 4096
    iterations: at a constant learning rate one call of 4100 iterations has
    the same loss history, bit for bit, as calls of 4096 and 4 (Adam's step
    count carries over between calls), and its first loss is the one
    loss_gradients reports."""
    widths = (4, 24, 8)
    x, t = data(widths, 40, 6)
    a, b = MLP(widths, seed=2), MLP(widths, seed=2)
    loss0, _ = b.loss_gradients(x, t)
    h = a.train(x, t, 4100, 1e-3, 1e-3)
    h1 = b.train(x, t, 4096, 1e-3, 1e-3)
    h2 = b.train(x, t, 4, 1e-3, 1e-3)
    assert h[0] == loss0
    assert np.array_equal(h, np.concatenate([h1, h2]))
    assert np.array_equal(a.get_parameters(), b.get_parameters())
    a.close()
    b.close()
