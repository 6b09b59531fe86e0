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
    """Parameter blocks; float64 parameters stay float64 (the exact path), others become float32."""
    dt = np.float64 if np.asarray(p).dtype == np.float64 else np.float32
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=dt).reshape(shape)
        o += n
    return out


def pack(d, dtype=np.float32):
    return np.concatenate([np.asarray(d[name], dtype=dtype).reshape(-1) for name, _ in SIZES])


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


def _rounding(emulate_bf16):
    """Operand rounding: bfloat16 (the device's matrix-core operands) or none."""
    return bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))


def forward(p, x, emulate_bf16=True):
    """x: (B, 6) -> y (B, 2) and the cache for backward (float64 math).  With
    emulate_bf16=False nothing is rounded: float64 parameters and inputs give
    the exact network."""
    P = unpack(p)
    q = _rounding(emulate_bf16)
    c32 = (lambda a: a.astype(np.float32)) if emulate_bf16 else (lambda a: a)  # noqa: E731
    xb = q(x).astype(np.float64)
    z1 = xb @ q(P["W1"]).astype(np.float64).T + P["b1"]
    h1 = q(c32(gelu(z1))).astype(np.float64)
    z2 = h1 @ q(P["W2"]).astype(np.float64).T + P["b2"]
    h2 = q(c32(gelu(z2))).astype(np.float64)
    y = h2 @ q(P["W3"]).astype(np.float64).T + P["b3"]
    return y, (xb, z1, h1, z2, h2)


def loss_and_grad(p, x, t, emulate_bf16=True):
    """Mean-squared error over the batch (sum over the 2 outputs) and its gradient.

    Mirrors the device kernel: every matrix-product operand is rounded to
    bfloat16 (weights, x, h1, h2, dy, dz2, dz1), products accumulate in FP32
    (here FP64), element-wise factors and bias gradients stay unrounded."""
    P = unpack(p)
    q = _rounding(emulate_bf16)
    c32 = (lambda a: a.astype(np.float32)) if emulate_bf16 else (lambda a: a)  # noqa: E731
    f64 = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731
    y, (xb, z1, h1, z2, h2) = forward(p, x, emulate_bf16)
    B = x.shape[0]
    e = y - t
    loss = float(np.sum(e * e) / B)
    dy = 2.0 * e / B
    dyb = f64(q(c32(dy)))
    g = {}
    g["W3"] = dyb.T @ h2
    g["b3"] = dy.sum(axis=0)
    dh2 = dyb @ f64(q(P["W3"]))
    dz2 = dh2 * gelu_grad(z2)
    dz2b = f64(q(c32(dz2)))
    g["W2"] = dz2b.T @ h1
    g["b2"] = dz2.sum(axis=0)
    dh1 = dz2b @ f64(q(P["W2"]))
    dz1 = dh1 * gelu_grad(z1)
    dz1b = f64(q(c32(dz1)))
    g["W1"] = dz1b.T @ xb
    g["b1"] = dz1.sum(axis=0)
    out_dt = np.float64 if np.asarray(p).dtype == np.float64 else np.float32
    return loss, pack(g, out_dt)


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in float32, as the device."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return (p - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32), m, v
