"""Write the exact direct-sum alm of the C1 workload (test fixture).

    python tests/golden/make_exact_fixtures.py

C1: Nl=256, 1e5 points uniform on the sphere, rng = numpy default_rng(1):
theta = arccos(U[-1, 1]), phi = U[0, 2 pi), w = U[0.5, 1.5] (drawn in that
order).  exact_c1.npz holds the generator's seed and size and
a_lm = sum_i w_i Y*_lm(theta_i, phi_i) from oracle.exact.exact_alm (FP64,
no spline), about 80 s on one core; the inputs are regenerated from the seed.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import exact  # noqa: E402


def c1_points(seed=1, n=100_000):
    rng = np.random.default_rng(seed)
    t = np.arccos(rng.uniform(-1, 1, n))
    p = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    return t, p, w


def main():
    t, p, w = c1_points()
    alm = exact.exact_alm(256, t, p, w)
    np.savez_compressed(os.path.join(HERE, "exact_c1.npz"), Nl=256, seed=1, n=len(t), alm=alm)
    print("wrote", os.path.join(HERE, "exact_c1.npz"))


if __name__ == "__main__":
    main()
