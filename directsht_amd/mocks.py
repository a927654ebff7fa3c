"""Synthetic point sets for the benchmark workloads (no healpy).

The reference draws mock catalogues with ``hp.synfast`` plus Poisson sampling
(sht/mocks.py:55-89).  healpy is not available here, so the Gaussian random
field is synthesized directly: a_lm ~ N(0, C_l) (a_l0 real), evaluated on a
(theta, phi) grid as g = sum_l a_l0 Y_l0 + 2 Re sum_{m>0} a_lm Y_lm e^{i m phi}
(normalized Legendre recurrence per ring, FFT over phi), rescaled to the
requested variance.  Points are then drawn by rejection with density
proportional to exp(g - sigma^2/2) (bilinear lookup in the grid), on the GPU
with torch's seeded generator.
"""
import functools
import math

import numpy as np
import torch


@functools.lru_cache(maxsize=4)
def _ylm_rings_cached(lmax, ntheta):
    theta = (np.arange(ntheta) + 0.5) * math.pi / ntheta
    Y = _ylm_rings(lmax, theta)
    Y.setflags(write=False)
    return Y


def _ylm_rings(lmax, theta):
    """Fully normalized Y_lm(theta, 0), l <= lmax, 0 <= m <= l: [m][l][ring]."""
    x, s = np.cos(theta), np.sin(theta)
    L = lmax + 1
    out = np.zeros((L, L, theta.size))
    pmm = np.full(theta.size, 1.0 / math.sqrt(4 * math.pi))
    for m in range(L):
        if m > 0:
            pmm = -math.sqrt((2 * m + 1) / (2.0 * m)) * s * pmm
        out[m, m] = pmm
        if m + 1 < L:
            p1 = math.sqrt(2 * m + 3.0) * x * pmm
            out[m, m + 1] = p1
            p2 = pmm
            for ell in range(m + 2, L):
                a = math.sqrt((4.0 * ell * ell - 1) / (ell * ell - m * m))
                b = math.sqrt(((ell - 1.0) ** 2 - m * m) / (4.0 * (ell - 1) ** 2 - 1))
                p = a * (x * p1 - b * p2)
                out[m, ell] = p
                p2, p1 = p1, p
    return out


def grf_grid(lmax=256, ntheta=512, nphi=1024, sigma=0.5, seed=0,
             cl=lambda ell: (ell + 10.0) ** -2.0):
    """Gaussian random field on a (theta, phi) grid, variance sigma^2.

    C_l = cl(l) for 1 <= l <= lmax, C_0 = 0.  theta rings at cell centres of
    [0, pi], phi samples 2 pi j / nphi.  Returns (theta_rings, g[ntheta, nphi]).
    """
    rng = np.random.default_rng(seed)
    theta = (np.arange(ntheta) + 0.5) * math.pi / ntheta
    ells = np.arange(lmax + 1)
    c = np.array([0.0] + [cl(l) for l in ells[1:]])
    a = np.zeros((lmax + 1, lmax + 1), dtype=np.complex128)  # [m][l]
    for m in range(lmax + 1):
        n = lmax + 1 - m
        if m == 0:
            a[0, :] = rng.normal(size=n) * np.sqrt(c)
        else:
            a[m, m:] = (rng.normal(size=n) + 1j * rng.normal(size=n)) * np.sqrt(c[m:] / 2)
    Y = _ylm_rings_cached(lmax, ntheta)               # [m][l][ring]
    F = np.einsum("ml,mlr->rm", a, Y)                 # F_m(theta) per ring
    spec = np.zeros((ntheta, nphi // 2 + 1), dtype=np.complex128)
    spec[:, :lmax + 1] = F
    g = np.fft.irfft(spec, n=nphi, axis=1) * nphi     # F_0 + 2 Re sum_m F_m e^{i m phi}
    g -= g.mean()
    g *= sigma / g.std()
    return theta, g


def uniform_sphere(n, seed=0, device="cuda", wlo=0.5, whi=1.5):
    """theta = arccos(U[-1,1]), phi = U[0, 2pi), w = U[wlo, whi] (FP64 tensors)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    u = torch.rand(n, dtype=torch.float64, device=device, generator=gen)
    p = torch.rand(n, dtype=torch.float64, device=device, generator=gen) * (2 * math.pi)
    w = wlo + (whi - wlo) * torch.rand(n, dtype=torch.float64, device=device, generator=gen)
    return torch.acos(2 * u - 1), p, w


def _bilinear(g, theta_rings, th, ph):
    """g sampled at (th, ph) by bilinear interpolation on the ring grid."""
    ntheta, nphi = g.shape
    dth = math.pi / ntheta
    ft = (th / dth - 0.5).clamp(0, ntheta - 1 - 1e-9)
    i0 = ft.floor().long()
    i1 = (i0 + 1).clamp(max=ntheta - 1)
    at = ft - i0
    fp = ph / (2 * math.pi) * nphi
    j0 = fp.floor().long() % nphi
    j1 = (j0 + 1) % nphi
    ap = fp - fp.floor()
    return ((1 - at) * ((1 - ap) * g[i0, j0] + ap * g[i0, j1])
            + at * ((1 - ap) * g[i1, j0] + ap * g[i1, j1]))


def lognormal_cap_points(n, seed=2, cmin=0.3, cmax=1.0, sigma=0.5, lmax=256, ntheta=512,
                         nphi=1024, device="cuda"):
    """n points in the cap cos(theta) in [cmin, cmax] with density
    proportional to exp(g - sigma^2/2), g a GRF with C_l ~ (l+10)^-2.
    Returns FP64 tensors (theta, phi, w = 1)."""
    theta_r, g = grf_grid(lmax, ntheta, nphi, sigma, seed)
    gt = torch.from_numpy(g).to(device)
    dens = torch.exp(gt - sigma ** 2 / 2)
    dmax = float(dens.max())
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    out_t, out_p, have = [], [], 0
    while have < n:
        m = max(1 << 20, int(1.2 * (n - have) * dmax))
        c = cmin + (cmax - cmin) * torch.rand(m, dtype=torch.float64, device=device, generator=gen)
        th = torch.acos(c)
        ph = torch.rand(m, dtype=torch.float64, device=device, generator=gen) * (2 * math.pi)
        u = torch.rand(m, dtype=torch.float64, device=device, generator=gen)
        acc = u * dmax < torch.exp(_bilinear(gt, theta_r, th, ph) - sigma ** 2 / 2)
        out_t.append(th[acc])
        out_p.append(ph[acc])
        have += int(acc.sum())
    t = torch.cat(out_t)[:n].contiguous()
    p = torch.cat(out_p)[:n].contiguous()
    return t, p, torch.ones_like(t)


def lognormal_weights(theta, phi, seed, sigma=0.5, lmax=256, ntheta=512, nphi=1024):
    """exp(g_seed - sigma^2/2) at the points, for many-weight-set workloads."""
    theta_r, g = grf_grid(lmax, ntheta, nphi, sigma, seed)
    gt = torch.from_numpy(g).to(theta.device)
    return torch.exp(_bilinear(gt, theta_r, theta, phi) - sigma ** 2 / 2)


def lognormal_weight_sets(theta, phi, nsets, seed0=300, sigma=0.5, lmax=256, ntheta=512, nphi=1024):
    """(nsets, N) weights: set 0 constant 1, set s > 0 = exp(g_s - sigma^2/2)
    with an independent field g_s (seed seed0 + s) -- the many-weight-set
    workload."""
    W = torch.empty((nsets, theta.numel()), dtype=torch.float64, device=theta.device)
    W[0] = 1.0
    for s in range(1, nsets):
        W[s] = lognormal_weights(theta, phi, seed0 + s, sigma, lmax, ntheta, nphi)
    return W
