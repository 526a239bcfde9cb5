"""CPU FP64 restatement of the surrogate network (numpy).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the
device network (nsm_mlp_* of the C ABI); the product never imports it.  The
network has no counterpart in the reference (NimbleSM has no learned
surrogate); what pins this restatement is calculus: its analytic gradients
agree with central differences of its own loss (tests/test_mlp_oracle.py),
and planted mistakes (a wrong activation derivative, a dropped bias
gradient, a transposed weight) fail that check.

Network: widths d_0 .. d_L, Z_l = H_{l-1} W_l^T + b_l, H_l = gelu(Z_l) for
l < L, output Y = Z_L; gelu(x) = x (1 + erf(x / sqrt 2)) / 2.  Loss: mean of
(Y - T)^2 over batch and outputs.  Parameters flat in the order W_1 [d_1][d_0]
(row-major), b_1, W_2, b_2, ...

With ``bf16=True`` the forward and backward passes round where the device
rounds to BF16 (inputs, weights, hidden activations, and the backward
deltas entering the products), so the restatement follows the device's
arithmetic up to its FP32 accumulation.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

SQRT1_2 = 1.0 / math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def to_bf16(x):
    """Round to FP32, then to BF16 (round to nearest even); returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    b = f.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x * SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x * SQRT1_2)) + x * INV_SQRT_2PI * np.exp(-0.5 * x * x)


def layout(widths):
    """[(w_offset, b_offset)] per layer and the parameter count."""
    off, out = 0, []
    for l in range(1, len(widths)):
        w = off
        off += widths[l] * widths[l - 1]
        out.append((w, off))
        off += widths[l]
    return out, off


def unpack(params, widths):
    lay, n = layout(widths)
    p = np.asarray(params, np.float64)
    assert p.size == n
    Ws = [p[w:w + widths[l + 1] * widths[l]].reshape(widths[l + 1], widths[l]) for l, (w, _) in enumerate(lay)]
    bs = [p[b:b + widths[l + 1]] for l, (_, b) in enumerate(lay)]
    return Ws, bs


def forward(params, widths, x, bf16=False):
    """Output Y and the per-layer (H_{l-1}, Z_l) of a forward pass."""
    Ws, bs = unpack(params, widths)
    rnd = to_bf16 if bf16 else (lambda v: np.asarray(v, np.float64))
    H = rnd(x)
    cache = []
    L = len(Ws)
    for l in range(L):
        Z = H @ rnd(Ws[l]).T + (bs[l].astype(np.float32).astype(np.float64) if bf16 else bs[l])
        cache.append((H, Z))
        if l < L - 1:
            H = rnd(gelu(Z))
    return Z, cache


def loss_and_grad(params, widths, x, t, bf16=False, act_grad=gelu_grad):
    """Mean squared error and its gradient in parameter order."""
    Ws, _ = unpack(params, widths)
    rnd = to_bf16 if bf16 else (lambda v: np.asarray(v, np.float64))
    Y, cache = forward(params, widths, x, bf16)
    n = Y.size
    E = Y - np.asarray(t, np.float64)
    loss = float(np.sum(E * E) / n)
    G = 2.0 * E / n
    lay, total = layout(widths)
    grad = np.zeros(total)
    for l in range(len(Ws) - 1, -1, -1):
        H, _ = cache[l]
        D = rnd(G)
        w, b = lay[l]
        grad[w:w + Ws[l].size] = (D.T @ H).ravel()
        grad[b:b + Ws[l].shape[0]] = G.sum(axis=0)
        if l > 0:
            G = (D @ rnd(Ws[l])) * act_grad(cache[l - 1][1])
    return loss, grad


def init_params(widths, seed):
    """The device's initialisation: weights uniform in +-sqrt(6/(fan_in+fan_out))
    from SplitMix64(seed) (53-bit uniforms), zero biases, rounded to FP32."""
    lay, n = layout(widths)
    p = np.zeros(n)
    s = seed & 0xFFFFFFFFFFFFFFFF
    mask = 0xFFFFFFFFFFFFFFFF

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        return z ^ (z >> 31)

    for l, (w, _) in enumerate(lay):
        a = math.sqrt(6.0 / (widths[l] + widths[l + 1]))
        for i in range(widths[l + 1] * widths[l]):
            u = (nxt() >> 11) * (1.0 / 9007199254740992.0)
            p[w + i] = (2.0 * u - 1.0) * a
    return p.astype(np.float32).astype(np.float64)
