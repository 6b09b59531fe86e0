"""Writes the small golden fixtures under tests/golden/ from the CPU oracle.

The reference (Julia) cannot run in this environment and ships no fixtures
for this path, so these vectors come from the oracle (oracle/collop_oracle.cpp),
whose functions are pinned independently by tests/test_oracle_*.py.  The
pair-sum fixture also carries a long-double brute-force result computed here
with numpy, as an oracle-independent anchor.

Run: python tests/golden/make_golden.py   (rewrites golden_*.npz)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "collisionoperators.jl_amd"))
sys.path.insert(0, os.path.dirname(HERE))

import collop  # noqa: E402
import oracle as O  # noqa: E402
from test_oracle_collision import brute_longdouble  # noqa: E402


def uniform(lo, hi, n):
    return np.array([lo + (hi - lo) * i / n for i in range(n + 1)])


def main():
    # pair sum: 96 markers, one outside the box, one coincident pair
    rng = np.random.default_rng(2024)
    v = O.colmajor(rng.normal(0, 1.5, (96, 2)))
    v[5] = (6.5, 0.0)
    v[7] = v[8]
    G = O.colmajor(rng.normal(0, 1, (96, 2)))
    w = rng.uniform(0.5, 1.5, 96)
    w /= w.sum()
    box = np.array([-6.0, 6.0, -6.0, 6.0])
    np.savez(os.path.join(HERE, "golden_pair.npz"), v=v, w=w, G=G, box=box,
             dot_v=O.collision(box, v, w, G), dot_v_ld=brute_longdouble(v, w, G, box).astype(np.float64))

    # projection chain on a 10 x 8 mesh
    bp1, bp2 = uniform(-4, 4, 10), uniform(-3, 3, 8)
    m = O.Mesh(bp1, bp2)
    v, w = collop.markers("gaussian", 700, 11, sigma1=1.3, sigma2=1.0)
    fc = m.l2_project(v, w)
    r = m.compute_r(fc)
    L = m.mass_solve(r)
    np.savez(os.path.join(HERE, "golden_fem.npz"), bp1=bp1, bp2=bp2, v=v, w=w, rhs=m.l2_rhs(v, w), fc=fc, r=r, L=L,
             S=m.entropy(fc), G=m.compute_G(v, L), g=m.loggrad(v, fc))

    # one implicit step, well conditioned (dt = 0.005)
    bp = uniform(-6, 6, 12)
    m = O.Mesh(bp, bp)
    v, w = collop.markers("bimodal", 300, 5, sigma1=0.8, sigma2=0.8, peak=2.0)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(0.005, m=5)
    v1, fc1, st = m.time_step(P, v, w, fc, S0)
    np.savez(os.path.join(HERE, "golden_step.npz"), bp=bp, v=v, w=w, dt=0.005, v_new=v1, fc_new=fc1,
             iters=st.iters, entropy=st.entropy, energy=st.energy)
    print("wrote golden_pair.npz golden_fem.npz golden_step.npz")


if __name__ == "__main__":
    main()
