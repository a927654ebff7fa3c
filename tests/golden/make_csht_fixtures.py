"""Write golden vectors from the reference's own C code (csht/sht_helper.c).

Run in the development container only, where the reference checkout exists:

    make -C oracle ref
    OMP_NUM_THREADS=1 python tests/golden/make_csht_fixtures.py

OMP_NUM_THREADS=1 is required: the contraction loop of do_transform
(csht/sht_helper.c:258-262) races on its thread index with more threads.
Only the resulting .npz data files are committed.

Fixtures
  csht_tables_small.npz  full make_table output, Nl=24, Nx=40, xmax=0.75
  csht_tables_rows.npz   selected rows (the sht/test_sht.py:44-45 grid plus
                         high-l rows) of make_table, Nl=500, Nx=1024, every
                         8th node
  csht_alm_equ.npz       do_transform alm via the csht/sht.py:78-140 driver:
                         2000 seeded equatorial points, Nl=64, Nx=128, and the
                         deterministic point set of csht/test_alm.py:66-77
                         (cos theta clipped to |cos theta| < 0.75), Nl=48, Nx=100 (chosen so that
                         no point lies on a node: csht bins by a scan
                         (csht/sht_helper.c:199-203) but interpolates with
                         (int)(ax/dx) (:208), which disagree at nodes)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import cref  # noqa: E402


def main():
    assert os.environ.get("OMP_NUM_THREADS") == "1", "run with OMP_NUM_THREADS=1"
    assert cref.available(), "build oracle/_ref first (make -C oracle ref)"
    # (a) full small tables
    c = cref.CRef(24, 40, 0.75)
    Yv, Yd = c.tables()
    np.savez_compressed(os.path.join(HERE, "csht_tables_small.npz"), Nl=24, Nx=40, xmax=0.75,
                        Yv=Yv, Yd=Yd)
    # (b) selected rows of a large table
    Nl, Nx = 500, 1024
    c = cref.CRef(Nl, Nx, 0.75)
    Yv, Yd = c.tables()
    lm = [(l, m) for l in (1, 10, 25, 40) for m in (0, 5, 15, 25, 40) if m <= l]
    lm += [(2, 0), (10, 10), (100, 0), (100, 50), (300, 150), (499, 0), (499, 100),
           (499, 250), (499, 499), (498, 497)]
    lm = sorted(set(lm))
    idx = np.array([(m * (2 * Nl - 1 - m)) // 2 + l for l, m in lm])
    cols = np.arange(0, Nx, 8)
    np.savez_compressed(os.path.join(HERE, "csht_tables_rows.npz"), Nl=Nl, Nx=Nx, xmax=0.75,
                        lm=np.array(lm), cols=cols, Yv=Yv[idx][:, cols], Yd=Yd[idx][:, cols])
    # (c) alm for equatorial points
    rng = np.random.default_rng(20240611)
    N = 2000
    ct = rng.uniform(-0.74, 0.74, N)
    theta = np.arccos(ct)
    phi = rng.uniform(0, 2 * np.pi, N)
    wt = rng.uniform(0.5, 1.5, N)
    c = cref.CRef(64, 128, 0.75)
    alm_rand = c(theta, phi, wt)
    cdet = np.concatenate([np.linspace(-0.7, -0.3, 500), np.linspace(0.2, 0.7, 500)])
    theta_d = np.arccos(cdet)
    phi_d = np.linspace(0.0, 2 * np.pi, 1000)
    wt_d = np.ones_like(theta_d)
    c = cref.CRef(48, 100, 0.75)
    alm_det = c(theta_d, phi_d, wt_d)
    np.savez_compressed(os.path.join(HERE, "csht_alm_equ.npz"),
                        Nl_rand=64, Nx_rand=128, theta=theta, phi=phi, wt=wt, alm_rand=alm_rand,
                        Nl_det=48, Nx_det=100, theta_det=theta_d, phi_det=phi_d, wt_det=wt_d,
                        alm_det=alm_det, xmax=0.75)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
