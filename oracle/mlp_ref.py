"""numpy reference of the per-marker warm-start MLP (TEST INFRASTRUCTURE ONLY).

The reference (junyixu/CollisionOperators.jl) has only the hook site of its
neural warm start (main.jl:164-166, 252-254; the model file warmstart_nn.jl is
absent, Parameters.jl:108-117), so there is no reference implementation to
restate.  This module defines the network that libcollop.so implements on the
GPU and is the checker of its forward and backward kernels:

    x in R^6 (normalised features)  ->  6 -> 64 -> 64 -> 2 with GELU
    z1 = W1 x + b1, h1 = gelu(z1); z2 = W2 h1 + b2, h2 = gelu(z2); y = W3 h2 + b3

GELU is the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).  The
GPU rounds weights and the MFMA operands (x, h1, h2 and the back-propagated
signals) to bfloat16 and accumulates in FP32; `bf16()` reproduces that
rounding (round to nearest even) so the comparison isolates accumulation-order
differences.  The per-marker output y is a velocity correction: the midpoint
velocity is predicted as v_euler + y * out_scale.  Parameter vector layout (float32, 4738 entries):
W1[64,6] b1[64] W2[64,64] b2[64] W3[2,64] b3[2], row-major.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 6, 64, 2
SIZES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (OUT, HID)),
         ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)
C0 = np.sqrt(2.0 / np.pi)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float32).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float32).reshape(-1) for name, _ in SIZES])


def bf16(a):
    """Round float32 to bfloat16 (nearest even) and back."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(z):
    z = z.astype(np.float64)
    t = np.tanh(C0 * (z + 0.044715 * z ** 3))
    return 0.5 * z * (1 + t)


def gelu_grad(z):
    z = z.astype(np.float64)
    u = C0 * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * C0 * (1 + 3 * 0.044715 * z * z)


def init(seed: int = 0):
    """Glorot-uniform weights, zero biases (same recipe as the device init)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            d[name] = rng.uniform(-lim, lim, shape).astype(np.float32)
        else:
            d[name] = np.zeros(shape, np.float32)
    return pack(d)


def forward(p, x, emulate_bf16=True):
    """x: (B, 6) float32 -> y (B, 2) and the cache for backward (float64 math)."""
    P = unpack(p)
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float32))
    xb = q(x).astype(np.float64)
    z1 = xb @ q(P["W1"]).astype(np.float64).T + P["b1"]
    h1 = q(gelu(z1).astype(np.float32)).astype(np.float64)
    z2 = h1 @ q(P["W2"]).astype(np.float64).T + P["b2"]
    h2 = q(gelu(z2).astype(np.float32)).astype(np.float64)
    y = h2 @ q(P["W3"]).astype(np.float64).T + P["b3"]
    return y, (xb, z1, h1, z2, h2)


def loss_and_grad(p, x, t, emulate_bf16=True):
    """Mean-squared error over the batch (sum over the 2 outputs) and its gradient.

    Mirrors the device kernel: every matrix-product operand is rounded to
    bfloat16 (weights, x, h1, h2, dy, dz2, dz1), products accumulate in FP32
    (here FP64), element-wise factors and bias gradients stay unrounded."""
    P = unpack(p)
    q = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float32))
    f64 = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731
    y, (xb, z1, h1, z2, h2) = forward(p, x, emulate_bf16)
    B = x.shape[0]
    e = y - t
    loss = float(np.sum(e * e) / B)
    dy = 2.0 * e / B
    dyb = f64(q(dy.astype(np.float32)))
    g = {}
    g["W3"] = dyb.T @ h2
    g["b3"] = dy.sum(axis=0)
    dh2 = dyb @ f64(q(P["W3"]))
    dz2 = dh2 * gelu_grad(z2)
    dz2b = f64(q(dz2.astype(np.float32)))
    g["W2"] = dz2b.T @ h1
    g["b2"] = dz2.sum(axis=0)
    dh1 = dz2b @ f64(q(P["W2"]))
    dz1 = dh1 * gelu_grad(z1)
    dz1b = f64(q(dz1.astype(np.float32)))
    g["W1"] = dz1b.T @ xb
    g["b1"] = dz1.sum(axis=0)
    return loss, pack({k: v.astype(np.float32) for k, v in g.items()})


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return (p - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32), m, v
