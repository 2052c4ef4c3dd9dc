"""Pins of the radiance-cache MLP oracle (oracle/mlp.py): its exact float64 back-propagation
against PyTorch autograd and finite differences, and its input encoding against closed forms."""
import math

import numpy as np
import pytest

from oracle import mlp


def _batch(rng, n):
    raw = np.zeros((n, 16), dtype=np.float32)
    raw[:, 0:3] = rng.uniform(0, 10, size=(n, 3))
    d = rng.normal(size=(n, 3))
    raw[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    raw[:, 6:9] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    raw[:, 9] = rng.uniform(0.05, 1, n)
    raw[:, 10:13] = rng.uniform(0, 1, size=(n, 3))
    raw[:, 13] = rng.uniform(0, 1, n) < 0.3
    target = rng.uniform(0, 3, size=(n, 3))
    return raw, target


def test_param_count():
    assert mlp.param_count() == 96 * 64 + 64 + 4 * (64 * 64 + 64) + 64 * 3 + 3 == 23043


def test_encoding_closed_forms():
    raw = np.zeros((2, 16), dtype=np.float32)
    raw[1, 0:3] = 5.0                      # the box centre: p = 0.5
    raw[:, 3:6] = [0.0, 0.0, 1.0]
    x = mlp.encode(raw, [0, 0, 0], [10, 10, 10])
    assert np.allclose(x[0, 0:72:2], 0.0) and np.allclose(x[0, 1:72:2], 1.0)      # p = 0: sin 0, cos 1
    for f in range(12):          # p = 0.5: sin(2^f pi / 2), up to the binary32 pi times 2^f
        assert abs(x[1, 2 * f] - math.sin(2 ** f * math.pi / 2)) < 1e-6 + 2 ** f * 1e-7
    assert np.allclose(x[:, 84:87], 0.0) and np.all(x[:, 92:] == 0.0)
    assert abs(x[0, 72 + 8] - 0.0) < 1e-6 and abs(x[0, 72 + 9] + 1.0) < 1e-6     # w_z = 1: sin pi, cos pi


def test_backprop_equals_autograd():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    raw, target = _batch(rng, 37)
    x = mlp.encode(raw, [0, 0, 0], [10, 10, 10])
    params = mlp.init_params(1).astype(np.float64)
    params += rng.normal(scale=0.01, size=params.shape)          # nonzero biases too
    loss, grad, _ = mlp.loss_and_grad(params, x, target, bf=False)
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(x, dtype=torch.float64)
    off = 0
    for L, (i, o) in enumerate(mlp.DIMS):
        w = p[off:off + i * o].reshape(o, i); off += i * o
        b = p[off:off + o]; off += o
        h = h @ w.T + b
        if L < mlp.LAYERS - 1:
            h = torch.relu(h)
    t = torch.tensor(target, dtype=torch.float64)
    tl = ((h - t) ** 2 / (h.detach() ** 2 + 0.01)).mean()
    tl.backward()
    assert abs(loss - tl.item()) < 1e-12
    np.testing.assert_allclose(grad, p.grad.numpy(), rtol=1e-9, atol=1e-14)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    raw, target = _batch(rng, 16)
    x = mlp.encode(raw, [0, 0, 0], [10, 10, 10])
    params = mlp.init_params(2).astype(np.float64)
    _, grad, y0 = mlp.loss_and_grad(params, x, target, bf=False)
    den = y0 * y0 + 0.01                      # the loss holds the denominator fixed

    def loss_at(p):
        y, _, _ = mlp.forward(p, x, bf=False)
        return np.mean((y - target) ** 2 / den)

    for k in rng.choice(len(params), size=40, replace=False):
        e = np.zeros_like(params)
        e[k] = 1e-6
        fd = (loss_at(params + e) - loss_at(params - e)) / 2e-6
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9


def test_fp16_rounding_closed_forms():
    """fp16() is IEEE binary16 round-to-nearest-even (ties to the even significand)."""
    one = 1.0
    assert mlp.fp16(one + 2.0 ** -11) == one                     # tie -> even (1.0)
    assert mlp.fp16(one + 3 * 2.0 ** -11) == one + 2.0 ** -9     # tie -> even (next up is odd)
    assert mlp.fp16(65504.0) == 65504.0 and np.isinf(mlp.fp16(65520.0))
    assert mlp.fp16(2.0 ** -24) == 2.0 ** -24 and mlp.fp16(2.0 ** -26) == 0.0


def test_loss_scale_is_divided_back_out():
    """Without rounding, back-propagating scale * dL/dy and dividing by scale is the identity."""
    rng = np.random.default_rng(3)
    x = rng.normal(size=(40, mlp.IN))
    t = rng.uniform(0, 2, size=(40, mlp.OUT))
    p = mlp.init_params(2).astype(np.float64)
    l1, g1, _ = mlp.loss_and_grad(p, x, t, bf=False)
    l2, g2, _ = mlp.loss_and_grad(p, x, t, bf=False, scale=1024.0)
    assert l1 == l2
    np.testing.assert_array_equal(g1, g2)
    # a nearly converged batch: dL/dy ~ 1e-8 is below binary16's smallest subnormal (6e-8), so
    # unscaled fp16 back-propagation loses the gradient and a scale of 1024 keeps it
    y, _, _ = mlp.forward(p, x, fmt="fp16")
    near = y * (1.0 + 1e-6) + 1e-7
    _, g16, _ = mlp.loss_and_grad(p, x, near, fmt="fp16", scale=1024.0)
    _, g16u, _ = mlp.loss_and_grad(p, x, near, fmt="fp16", scale=1.0)
    _, g64, _ = mlp.loss_and_grad(p, x, near, fmt="fp16", scale=2.0 ** 20)
    assert np.count_nonzero(g16u) < 0.5 * np.count_nonzero(g16)
    assert np.linalg.norm(g16 - g64) < 0.1 * np.linalg.norm(g64)


def test_adam_step_matches_torch_and_the_constant_gradient_closed_form():
    """oracle/mlp.py's adam_step against torch.optim.Adam (same betas and eps, float64) over six
    steps of random gradients, and against Adam's closed form for a constant gradient: every
    step moves each weight by lr * g / (|g| + eps'), i.e. by lr * sign(g) for |g| >> eps (a
    missing or swapped bias correction breaks this from step 2 on, a sign slip at once)."""
    import torch
    rng = np.random.default_rng(9)
    w0 = rng.normal(size=200)
    grads = [rng.normal(scale=10.0 ** rng.uniform(-3, 1), size=200) for _ in range(6)]
    p = torch.nn.Parameter(torch.tensor(w0, dtype=torch.float64))
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.99), eps=1e-8)
    w, m, v = w0.copy(), np.zeros(200), np.zeros(200)
    for t, g in enumerate(grads, start=1):
        p.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        w, m, v = mlp.adam_step(w, m, v, g, t, 1e-3)
        np.testing.assert_allclose(w, p.detach().numpy(), rtol=0, atol=1e-14)
    g = rng.choice([-1.0, 1.0], size=200) * rng.uniform(0.01, 5.0, size=200)
    w, m, v = w0.copy(), np.zeros(200), np.zeros(200)
    for t in range(1, 6):
        prev = w
        w, m, v = mlp.adam_step(w, m, v, g, t, 1e-3)
        np.testing.assert_allclose(w - prev, -1e-3 * np.sign(g), rtol=1e-5, atol=0)
