"""CPU restatement of the radiance-cache MLP (nbounce_amd/csrc/mlp.hip, cache.cpp) in float64.

TEST INFRASTRUCTURE ONLY (the checker of the GPU kernels).  `bf=True` rounds every value the
GPU feeds to the matrix cores to bfloat16 (round to nearest even) at the same places: the encoded
input, each hidden activation, the output gradient and each back-propagated gradient; the
rounding is treated as the identity in the backward pass, as on the GPU.  `fmt="fp16"` rounds to
IEEE binary16 instead, with the output gradient multiplied by the loss scale before rounding and
the weight gradients divided by it at the end, as the FP16 kernels do.  With `bf=False` and no
`fmt` it is the exact float64 network, which the tests check against PyTorch autograd.
"""
import math

import numpy as np

IN, HID, OUT, LAYERS = 96, 64, 3, 6
DIMS = [(IN, HID), (HID, HID), (HID, HID), (HID, HID), (HID, HID), (HID, OUT)]   # (in, out)


def param_count():
    return sum(i * o + o for i, o in DIMS)


def unpack(params):
    """Flat vector -> [(W (out x in), b)] in the order W0, b0, ..., W5, b5."""
    out, off = [], 0
    for i, o in DIMS:
        w = np.asarray(params[off:off + i * o], dtype=np.float64).reshape(o, i)
        off += i * o
        b = np.asarray(params[off:off + o], dtype=np.float64)
        off += o
        out.append((w, b))
    return out


def pack(grads):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in grads])


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def fp16(x):
    """Round to IEEE binary16 (nearest even) and return as float64."""
    with np.errstate(over="ignore"):
        return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float64)


def _rounding(bf, fmt):
    if fmt == "fp16":
        return fp16
    if fmt == "bf16" or (fmt is None and bf):
        return bf16
    return lambda v: np.asarray(v, np.float64)


def encode(raw, scene_min, scene_max):
    """The 96 input features of each raw record (see mlp.hip `feature`)."""
    raw = np.asarray(raw, dtype=np.float32)
    n = len(raw)
    x = np.zeros((n, IN), dtype=np.float64)
    inv = np.float32(1.0) / (np.asarray(scene_max, np.float32) - np.asarray(scene_min, np.float32))
    pi = np.float32(3.14159265359)
    for d in range(3):
        p = (raw[:, d] - np.float32(scene_min[d])) * inv[d]
        for f in range(12):
            a = (p * (pi * np.float32(2.0 ** f))).astype(np.float64)
            x[:, 24 * d + 2 * f] = np.sin(a)
            x[:, 24 * d + 2 * f + 1] = np.cos(a)
    for d in range(3):
        for f in range(2):
            a = (raw[:, 3 + d] * (pi * np.float32(2.0 ** f))).astype(np.float64)
            x[:, 72 + 4 * d + 2 * f] = np.sin(a)
            x[:, 72 + 4 * d + 2 * f + 1] = np.cos(a)
    x[:, 84:87] = raw[:, 6:9]
    x[:, 87] = raw[:, 9]
    x[:, 88:91] = raw[:, 10:13]
    x[:, 91] = raw[:, 13]
    return x


def forward(params, x, bf=True, fmt=None):
    """Returns (y, activations) with activations[0] = the (rounded) input."""
    r = _rounding(bf, fmt)
    layers = unpack(params)
    ws = [(r(w), b) for w, b in layers]
    h = [r(x)]
    for L in range(LAYERS - 1):
        w, b = ws[L]
        h.append(r(np.maximum(h[-1] @ w.T + b, 0.0)))
    w, b = ws[-1]
    y = h[-1] @ w.T + b
    return y, h, ws


def loss_and_grad(params, x, target, bf=True, fmt=None, scale=1.0):
    """Relative L2 loss mean((y - t)^2 / (y^2 + 0.01)) with y held constant in the denominator,
    and its gradient with respect to the flat parameter vector (back-propagated times `scale`,
    returned divided by it)."""
    r = _rounding(bf, fmt)
    y, h, ws = forward(params, x, bf, fmt)
    n = len(x)
    den = y * y + 0.01
    loss = float(np.mean((y - target) ** 2 / den))
    dz = r(2.0 * (y - target) / den / (3.0 * n) * scale)
    grads = [None] * LAYERS
    for L in range(LAYERS - 1, -1, -1):
        w, _ = ws[L]
        grads[L] = (dz.T @ h[L], dz.sum(axis=0))
        if L > 0:
            dz = r((dz @ w) * (h[L] > 0))
    return loss, pack(grads) / scale, y


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    ps = []
    for L, (i, o) in enumerate(DIMS):
        bound = math.sqrt(6.0 / i) * (0.1 if L == LAYERS - 1 else 1.0)
        ps.append(rng.uniform(-bound, bound, size=o * i))
        ps.append(np.zeros(o))
    return np.concatenate(ps).astype(np.float32)


def adam_step(w, m, v, g, t, lr, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in float64 on the flat parameters: step t
    (1-based) with gradient g; returns the new (w, m, v).  The update is
    lr * m_hat / (sqrt(v_hat) + eps) with m_hat = m / (1 - b1^t), v_hat = v / (1 - b2^t); the
    cache trains with b1 = 0.9, b2 = 0.99, eps = 1e-8 (cache.cpp)."""
    w, m, v, g = (np.asarray(a, dtype=np.float64) for a in (w, m, v, g))
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    m_hat = m / (1.0 - b1 ** t)
    v_hat = v / (1.0 - b2 ** t)
    return w - lr * m_hat / (np.sqrt(v_hat) + eps), m, v
