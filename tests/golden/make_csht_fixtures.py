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
  csht_c2_equ.npz        the headline configuration Nl=1000, Nx=2048: the
                         equatorial part (0.3 <= cos theta < 0.75) of 12000
                         points drawn by the lognormal-cap generator
                         (directsht_amd.mocks.lognormal_cap_points on the CPU,
                         its field synthesized by oracle.exact.synth_rings,
                         seed 2, w = 1), do_transform alm at every 37th
                         healpy index, and table rows at every 997th index
                         (every 16th node)
  csht_nl2100.npz        band limit above 2048: Nl=2100, Nx=40, 1500 seeded
                         equatorial points with weights U[0.5, 1.5]; alm at
                         every 101st index and table rows at every 4999th
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
    # (d) the headline configuration; the generator's field is synthesized
    # on the host by the exact oracle's ring synthesis instead of the device
    # kernel (same alm and grid), and the points are drawn by torch on the CPU
    import torch

    from directsht_amd import mocks
    from oracle import exact

    def synthesize_map_cpu(alm, Nl, x, nphi, device=None):
        F = torch.from_numpy(exact.synth_rings(np.asarray(alm), Nl, np.asarray(x)))
        spec = torch.zeros((F.shape[0], nphi // 2 + 1), dtype=torch.complex128)
        spec[:, :Nl] = F
        spec[:, 0] = spec[:, 0].real
        return torch.fft.irfft(spec, n=nphi, dim=1) * nphi

    mocks.synthesize_map = synthesize_map_cpu
    th, ph, _ = mocks.lognormal_cap_points(12000, seed=2, device="cpu")
    th, ph = th.numpy(), ph.numpy()
    equ = np.abs(np.cos(th)) < 0.75
    th, ph = th[equ], ph[equ]
    wt = np.ones_like(th)
    Nl, Nx = 1000, 2048
    c = cref.CRef(Nl, Nx, 0.75)
    alm = c(th, ph, wt)
    Yv, Yd = c.tables()
    aidx = np.arange(0, c.Nlm, 37)
    ridx = np.arange(0, c.Nlm, 997)
    cols = np.arange(0, Nx, 16)
    np.savez_compressed(os.path.join(HERE, "csht_c2_equ.npz"), Nl=Nl, Nx=Nx, xmax=0.75,
                        theta=th, phi=ph, wt=wt, aidx=aidx, alm=alm[aidx], ridx=ridx, cols=cols,
                        Yv=Yv[ridx][:, cols], Yd=Yd[ridx][:, cols])
    del c, Yv, Yd, torch
    # (e) a band limit above 2048
    Nl, Nx = 2100, 40
    rng = np.random.default_rng(2100)
    N = 1500
    theta = np.arccos(rng.uniform(-0.74, 0.74, N))
    phi = rng.uniform(0, 2 * np.pi, N)
    wt = rng.uniform(0.5, 1.5, N)
    c = cref.CRef(Nl, Nx, 0.75)
    alm = c(theta, phi, wt)
    Yv, Yd = c.tables()
    aidx = np.arange(0, c.Nlm, 101)
    ridx = np.arange(0, c.Nlm, 4999)
    np.savez_compressed(os.path.join(HERE, "csht_nl2100.npz"), Nl=Nl, Nx=Nx, xmax=0.75,
                        theta=theta, phi=phi, wt=wt, aidx=aidx, alm=alm[aidx], ridx=ridx,
                        Yv=Yv[ridx], Yd=Yd[ridx])
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
