"""Write tests/golden/golden_mlp_fd.npz: a central-difference gradient of the
warm-start MLP's batch loss, taken through the numpy restatement's forward
pass only (oracle/mlp_ref.py, exact float64 path, no bfloat16 rounding).

The fixture pins the restatement's analytic backward pass (loss_and_grad),
which the GPU gradient kernels are compared with: differencing the forward
pass checks the backward pass independently of its formulas.

Usage: python tests/golden/make_mlp_fd.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import mlp_ref as M  # noqa: E402


def main():
    rng = np.random.default_rng(2024)
    p = M.init(7).astype(np.float64)
    p[:384] *= 3.0  # first layer weights x3: activations leave GELU's linear range
    nb = M.NPARAM
    # small non-zero biases so every block of the gradient is exercised
    for name, sl in (("b1", slice(384, 448)), ("b2", slice(4544, 4608)), ("b3", slice(4736, 4738))):
        p[sl] = rng.normal(0.0, 0.1, sl.stop - sl.start)
    x = rng.normal(0.0, 1.0, (16, M.IN))
    t = rng.normal(0.0, 0.5, (16, M.OUT))

    def loss(q):
        y, _ = M.forward(q, x, emulate_bf16=False)
        return float(np.sum((y - t) ** 2) / x.shape[0])

    g = np.zeros(nb)
    h = 1e-5
    for k in range(nb):
        e = np.zeros(nb)
        e[k] = h
        # fourth-order central difference: error O(h^4)
        g[k] = (-loss(p + 2 * e) + 8 * loss(p + e) - 8 * loss(p - e) + loss(p - 2 * e)) / (12 * h)
    np.savez_compressed(os.path.join(HERE, "golden_mlp_fd.npz"), p=p, x=x, t=t, loss=loss(p), grad_fd=g, h=h)
    print(f"wrote golden_mlp_fd.npz: loss {loss(p):.6e}, |grad| {np.linalg.norm(g):.4e}")


if __name__ == "__main__":
    main()
