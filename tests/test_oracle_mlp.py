"""Pins of the warm-start MLP restatement (oracle/mlp_ref.py) that the GPU
forward, gradient and Adam kernels are checked against.

  * the analytic gradient against a committed fourth-order central-difference
    gradient of the forward pass (tests/golden/golden_mlp_fd.npz, written by
    tests/golden/make_mlp_fd.py);
  * GELU (tanh form): its published value at 1, the identity
    gelu(z) - gelu(-z) = z, and its derivative against differences;
  * Adam's first step in closed form, p - lr g / (|g| + eps);
  * bfloat16 round-to-nearest-even on exactly representable and tie inputs."""
import os

import numpy as np

import mlp_ref as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_mlp_fd.npz")


def test_gradient_matches_finite_differences():
    z = np.load(GOLD)
    loss, g = M.loss_and_grad(z["p"], z["x"], z["t"], emulate_bf16=False)
    assert abs(loss - float(z["loss"])) <= 1e-12 * abs(float(z["loss"]))
    fd = z["grad_fd"]
    err = np.max(np.abs(g - fd)) / np.max(np.abs(fd))
    print(f"analytic vs finite-difference gradient: {err:.2e}")
    assert err <= 1e-6
    # every parameter block carries gradient (the fixture exercises all of them)
    for sl in (slice(0, 384), slice(384, 448), slice(448, 4544), slice(4544, 4608), slice(4608, 4736),
               slice(4736, 4738)):
        assert np.max(np.abs(fd[sl])) > 1e-6


def test_gradient_fixture_rejects_a_wrong_backward():
    """A plausible slip (the GELU derivative without its second term) moves the
    gradient far outside the fixture's tolerance."""
    z = np.load(GOLD)
    saved = M.gelu_grad
    try:
        M.gelu_grad = lambda q: 0.5 * (1 + np.tanh(M.C0 * (q + 0.044715 * q ** 3)))
        _, g = M.loss_and_grad(z["p"], z["x"], z["t"], emulate_bf16=False)
    finally:
        M.gelu_grad = saved
    assert np.max(np.abs(g - z["grad_fd"])) / np.max(np.abs(z["grad_fd"])) > 1e-3


def test_gelu_known_values():
    assert abs(float(M.gelu(np.array([1.0]))[0]) - 0.8411919906) < 1e-9
    assert float(M.gelu(np.array([0.0]))[0]) == 0.0
    zz = np.linspace(-4, 4, 33)
    np.testing.assert_allclose(M.gelu(zz) - M.gelu(-zz), zz, atol=1e-14)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(zz), (M.gelu(zz + h) - M.gelu(zz - h)) / (2 * h), atol=1e-8)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(1)
    p = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    p1, m, v = M.adam(p, g, np.zeros(100), np.zeros(100), step=1, lr=1e-3)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-7)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-6)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1.0 + 2.0 ** -7], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3 * 2^-8 ties to 1 + 2^-6
    np.testing.assert_array_equal(M.bf16(x), np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 1.0 + 2.0 ** -7],
                                                       np.float32))
