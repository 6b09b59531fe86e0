"""Write tests/golden/stiff_bimodal_n8192.npz: the markers of a bimodal run
(N = 8192, two Maxwellians at (+-2, 0) with covariance 0.5 I, PCG64 seed 42,
32 x 32 cells on [-6, 6]^2) after 16 oracle steps of dt = 0.02 (Anderson,
m = 5, tol 1e-12).  By then close pairs have formed: the oracle's own
Anderson solve of the next step stagnates, so the state exercises the
stiff-cluster elimination.

Usage: python tests/golden/make_stiff_state.py   (about a minute on 8 cores)
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..", "..")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "collisionoperators.jl_amd"))
import oracle as O  # noqa: E402
from collop.ic import markers  # noqa: E402

N, STEPS, DT = 8192, 16, 0.02


def main():
    bp = np.array([-6 + 12 * i / 32 for i in range(33)])
    m = O.Mesh(bp, bp)
    v, w = markers("bimodal", N, 42, sigma1=math.sqrt(0.5), sigma2=math.sqrt(0.5), peak=2.0)
    v = O.colmajor(v)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(DT, m=5)
    for s in range(STEPS):
        v, fc, st = m.time_step(P, v, w, fc, S0)
        S0 = st.entropy
        print(f"step {s}: iters {st.iters} exit {st.exit_reason}", flush=True)
    np.savez_compressed(os.path.join(HERE, "stiff_bimodal_n8192.npz"), v=np.asarray(v), w=w, steps=STEPS, dt=DT)


if __name__ == "__main__":
    main()
