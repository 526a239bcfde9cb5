"""The CPU restatement of the surrogate network (oracle/mlp.py), pinned by
calculus: analytic gradients against central differences of its own loss,
planted mistakes caught, and the BF16 rounding against torch's conversion."""
import numpy as np
import pytest

from oracle import mlp as M

WIDTHS = (4, 7, 6, 3)


def setup(seed=0, batch=9):
    rng = np.random.default_rng(seed)
    _, n = M.layout(WIDTHS)
    p = rng.normal(0, 0.5, n)
    x = rng.normal(0, 1, (batch, WIDTHS[0]))
    t = rng.normal(0, 1, (batch, WIDTHS[-1]))
    return p, x, t


def central_differences(p, x, t, h=1e-6):
    g = np.zeros_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        g[i] = (M.loss_and_grad(p + e, WIDTHS, x, t)[0] - M.loss_and_grad(p - e, WIDTHS, x, t)[0]) / (2 * h)
    return g


def test_gradients_match_central_differences():
    p, x, t = setup()
    _, g = M.loss_and_grad(p, WIDTHS, x, t)
    fd = central_differences(p, x, t)
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) < 1e-7
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_gelu_closed_forms():
    assert M.gelu(0.0) == 0.0
    x = np.linspace(-6, 6, 101)
    # gelu(x) - gelu(-x) = x, and the derivative against differences of gelu
    np.testing.assert_allclose(M.gelu(x) - M.gelu(-x), x, rtol=0, atol=1e-14)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(x), (M.gelu(x + h) - M.gelu(x - h)) / (2 * h), rtol=0, atol=1e-8)


@pytest.mark.parametrize("mistake", ["activation", "bias", "transpose"])
def test_planted_mistakes_fail(mistake):
    p, x, t = setup(1)
    fd = central_differences(p, x, t)
    if mistake == "activation":
        _, g = M.loss_and_grad(p, WIDTHS, x, t, act_grad=M.gelu)
    else:
        _, g = M.loss_and_grad(p, WIDTHS, x, t)
        lay, _ = M.layout(WIDTHS)
        if mistake == "bias":
            w, b = lay[-1]
            g[b:b + WIDTHS[-1]] = 0.0
        else:  # the first layer's weight gradient stored transposed
            w, _ = lay[0]
            blk = g[w:w + WIDTHS[1] * WIDTHS[0]].reshape(WIDTHS[1], WIDTHS[0])
            g[w:w + blk.size] = blk.T.reshape(-1) if blk.shape[0] == blk.shape[1] else blk[::-1].reshape(-1)
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) > 1e-3


def test_bf16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(0, 1, 5000) * 10.0 ** rng.integers(-30, 30, 5000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 65504.0, 1e-40]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.to_bf16(x), ref)


def test_bf16_forward_close_to_fp64_forward():
    p, x, t = setup(3)
    y64, _ = M.forward(p, WIDTHS, x)
    y16, _ = M.forward(p, WIDTHS, x, bf16=True)
    err = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-5 < err < 3e-2  # rounding is there, and it is BF16-sized
