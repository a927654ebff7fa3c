"""Synthetic point sets: the reference's lognormal mock catalogues
(sht/mocks.py, LogNormalMocks) and the benchmark workloads, without healpy.

The reference draws a Gaussian random field with ``hp.synfast`` and Poisson
samples exp(g) per HEALPix pixel (sht/mocks.py:55-89).  Here the field is
synthesized on the device: a_lm ~ N(0, C_l) (a_l0 real), then
F_m(theta_r) = sum_l a_lm Ybar_lm(theta_r) on iso-latitude rings
(``dsht_synthesis_rings``, a HIP kernel in libdsht.so) and one inverse real
FFT per ring, g = F_0 + 2 Re sum_{m>0} F_m e^{i m phi}.  Points are drawn on
the device with torch's seeded generator: Poisson counts per pixel of an
equiangular ring grid (LogNormalMocks), or rejection sampling against a
bilinear lookup in the grid (the benchmark's cap workload).
"""
import math

import numpy as np
import torch


def synthesize_rings(alm, Nl, x, device=None, stream=None):
    """F[r, m] = sum_{l>=m} a_lm Ybar_lm(x_r) (complex128 CUDA tensor, shape
    (len(x), Nl)) for healpy-ordered alm, via dsht_synthesis_rings."""
    from . import _lib
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    a = alm if isinstance(alm, torch.Tensor) else torch.from_numpy(np.asarray(alm, np.complex128))
    a = a.to(device=device, dtype=torch.complex128).contiguous()
    xs = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.asarray(x, np.float64))
    xs = xs.to(device=device, dtype=torch.float64).contiguous()
    if a.numel() != Nl * (Nl + 1) // 2:
        raise ValueError(f"alm has {a.numel()} values, expected Nl(Nl+1)/2 = {Nl * (Nl + 1) // 2}")
    lib = _lib.load()
    with torch.cuda.device(device):
        F = torch.empty((xs.numel(), Nl), dtype=torch.complex128, device=device)
        st = stream if stream is not None else torch.cuda.current_stream(device)
        _lib.check(lib.dsht_synthesis_rings(Nl, a.data_ptr(), xs.numel(), xs.data_ptr(),
                                            F.data_ptr(), st.cuda_stream), "dsht_synthesis_rings")
    return F


def synthesize_map(alm, Nl, x, nphi, device=None):
    """Real field f(x_r, phi_j = 2 pi j / nphi) = Re F_0 + 2 Re sum_{m>0} F_m
    e^{i m phi_j} on the ring grid (float64 CUDA tensor (len(x), nphi)),
    nphi >= 2 Nl - 1 so that no order aliases."""
    if nphi < 2 * Nl - 1:
        raise ValueError("nphi must be at least 2 Nl - 1")
    F = synthesize_rings(alm, Nl, x, device)
    spec = torch.zeros((F.shape[0], nphi // 2 + 1), dtype=torch.complex128, device=F.device)
    spec[:, :Nl] = F
    spec[:, 0] = spec[:, 0].real
    return torch.fft.irfft(spec, n=nphi, dim=1) * nphi


def gaussian_alm(cl, gen):
    """a_lm ~ N(0, C_l) in healpy order (a_l0 real; Re, Im of a_{m>0} each of
    variance C_l/2) as a complex128 tensor, drawn with the torch generator."""
    cl = torch.as_tensor(np.asarray(cl, np.float64), device=gen.device)
    Nl = cl.numel()
    ell_h = np.concatenate([np.arange(m, Nl) for m in range(Nl)])
    ell = torch.from_numpy(ell_h).to(gen.device)
    m0 = torch.zeros(ell_h.size, dtype=torch.bool, device=gen.device)
    m0[:Nl] = True                                   # the m = 0 row comes first
    sd = torch.sqrt(cl[ell].clamp(min=0))
    re = torch.randn(ell.numel(), dtype=torch.float64, device=gen.device, generator=gen)
    im = torch.randn(ell.numel(), dtype=torch.float64, device=gen.device, generator=gen)
    s = torch.where(m0, sd, sd / math.sqrt(2.0))
    return torch.complex(re * s, torch.where(m0, torch.zeros_like(im), im * s))


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
    alm = np.concatenate([a[m, m:] for m in range(lmax + 1)])   # healpy order
    g = synthesize_map(alm, lmax + 1, np.cos(theta), nphi).cpu().numpy()
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


# ---------------------------------------------------------------------------
# The reference's lognormal mock catalogues (sht/mocks.py:10-127)
# ---------------------------------------------------------------------------
def _legendre_rows(lmax, x):
    """P_l(x) for l = 0..lmax, shape (lmax+1, len(x)), by Bonnet's recurrence
    (l+1) P_{l+1} = (2l+1) x P_l - l P_{l-1}."""
    x = np.asarray(x, dtype=np.float64)
    P = np.empty((lmax + 1, x.size))
    P[0] = 1.0
    if lmax >= 1:
        P[1] = x
    for l in range(1, lmax):
        P[l + 1] = ((2 * l + 1) * x * P[l] - l * P[l - 1]) / (l + 1)
    return P


def get_corrfunc_from_Cl(cls, xs):
    """w(x) = sum_l (2l+1) C_l P_l(x) / 4pi at x = cos(theta)
    (sht/mocks.py:129-143)."""
    cls = np.asarray(cls, dtype=np.float64)
    P = _legendre_rows(cls.size - 1, xs)
    return ((2 * np.arange(cls.size) + 1) * cls) @ P / (4 * np.pi)


def get_Cl_from_corrfunc(cf_at_xs, xs, weights, lmax):
    """C_l = 2pi sum_i w_i cf(x_i) P_l(x_i), l = 0..lmax, by Gauss-Legendre
    quadrature at nodes xs with weights (sht/mocks.py:145-163)."""
    P = _legendre_rows(lmax, xs)
    return 2 * np.pi * (P @ (np.asarray(weights) * np.asarray(cf_at_xs)))


def get_gauss_cl_from_ln_cl(ln_cl, gauss_order=1000):
    """C_l of the Gaussian field whose exponential has the lognormal C_l
    ln_cl: w_G = log(1 + w_LN) (sht/mocks.py:165-178)."""
    from numpy.polynomial.legendre import leggauss
    xs, weights = leggauss(gauss_order)
    ln_corr = get_corrfunc_from_Cl(ln_cl, xs)
    return get_Cl_from_corrfunc(np.log(ln_corr + 1), xs, weights, len(ln_cl) - 1)


class LogNormalMocks:
    def __init__(self, Npnt, nside=2048, lmax=1000, clg=None, cl_ln=None, norm=0.01, alpha=2.,
                 ell0=10., theta_range=(0, np.pi), phi_range=(0, 2 * np.pi), verbose=False,
                 device=None):
        """A lognormal catalogue of about Npnt points (sht/mocks.py:10-42).
        The Gaussian field has C_l = clg, or the C_l whose lognormal field has
        C_l = cl_ln, or norm (ell0/(l+ell0))^alpha for l < lmax.  The field is
        sampled on 3 nside rings x 4 nside longitudes (as many pixels as a
        HEALPix map of that nside, equiangular instead of equal-area)."""
        assert (cl_ln is None) or (clg is None), "Cannot specify both cl_ln and clg"
        self.Npnt = Npnt
        self.nside = nside
        self.verbose = verbose
        self.theta_range = theta_range
        self.phi_range = phi_range
        self.device = device
        if cl_ln is not None:
            self.clg = get_gauss_cl_from_ln_cl(cl_ln)
        elif clg is not None:
            self.clg = np.asarray(clg, dtype=np.float64)
        else:
            ell = np.arange(lmax)
            self.clg = norm * (ell0 / (ell + ell0)) ** alpha

    def __call__(self, seed=None, verbose=True):
        """(theta, phi, wt) of the catalogue inside theta_range x phi_range
        (sht/mocks.py:43-53); numpy arrays."""
        dev = self.device if self.device is not None else \
            torch.device("cuda", torch.cuda.current_device())
        self.gen = torch.Generator(device=dev)   # the reference's self.rng
        if seed is None:
            self.gen.seed()
        else:
            self.gen.manual_seed(int(seed))
        t, p, w = self.lognormal_catalog(verbose)
        keep = self.make_mask(t, p)
        return t[keep], p[keep], w[keep]

    def _grid(self):
        nth, nph = 3 * self.nside, 4 * self.nside
        nph = max(nph, 2 * self.clg.size - 1)   # no aliasing of the field's orders
        return nth, nph

    def lognormal_catalog(self, verbose=False):
        """(theta, phi, wt = 1) of Poisson(Npnt) points with intensity
        proportional to exp(g) (sht/mocks.py:55-88), drawn from the generator
        that __call__ seeds (successive calls continue its stream).  The intensity is
        constant over each pixel (value at its centre) and the points are
        uniform in its area, so any number of points may share a pixel."""
        dev = self.device if self.device is not None else \
            torch.device("cuda", torch.cuda.current_device())
        gen = getattr(self, "gen", None)
        if gen is None:  # called before __call__: a fresh, randomly seeded stream
            gen = self.gen = torch.Generator(device=dev)
            gen.seed()
        Nl = self.clg.size
        nth, nph = self._grid()
        edges = torch.linspace(0.0, math.pi, nth + 1, dtype=torch.float64, device=dev)
        xc = torch.cos(0.5 * (edges[:-1] + edges[1:]))
        alm = gaussian_alm(self.clg, gen)
        g = synthesize_map(alm, Nl, xc, nph, dev)
        if verbose:
            print("gmap in range ({:e},{:e})".format(float(g.min()), float(g.max())))
        # ring band areas (2 pi / nph) (cos t_r - cos t_{r+1})
        xe = torch.cos(edges)
        band = (xe[:-1] - xe[1:]).clamp(min=0)
        lam = torch.exp(g) * band[:, None]
        lam *= self.Npnt / float(lam.sum())
        counts = torch.poisson(lam.reshape(-1), generator=gen).to(torch.int64)
        pix = torch.repeat_interleave(torch.arange(counts.numel(), device=dev), counts)
        r, j = pix // nph, pix % nph
        u = torch.rand(pix.numel(), dtype=torch.float64, device=dev, generator=gen)
        v = torch.rand(pix.numel(), dtype=torch.float64, device=dev, generator=gen)
        x = xe[r] - u * (xe[r] - xe[r + 1])
        theta = torch.acos(x.clamp(-1.0, 1.0))
        phi = (j.to(torch.float64) + v) * (2 * math.pi / nph)
        if verbose:
            print(f"{pix.numel()} points on {nth}x{nph} pixels")
        t, p = theta.cpu().numpy(), phi.cpu().numpy()
        return t, p, np.ones_like(t)

    def make_mask(self, thetas, phis):
        """Indices of the points strictly inside theta_range x phi_range
        (sht/mocks.py:90-99)."""
        return np.nonzero((thetas > self.theta_range[0]) & (thetas < self.theta_range[1])
                          & (phis > self.phi_range[0]) & (phis < self.phi_range[1]))[0]

    def get_theory_Cl(self, lmax_out=None, shot_noise=True, gauss_order=1000):
        """C_l of the lognormal field, w_LN = exp(w_G) - 1 by Gauss-Legendre
        quadrature, plus the shot noise 4pi/Npnt (sht/mocks.py:101-127)."""
        from numpy.polynomial.legendre import leggauss
        if lmax_out is None:
            lmax_out = len(self.clg) - 1
        xs, weights = leggauss(gauss_order)
        ln_corr = np.exp(get_corrfunc_from_Cl(self.clg, xs)) - 1
        cls = get_Cl_from_corrfunc(ln_corr, xs, weights, lmax_out)
        return cls + (4 * np.pi / float(self.Npnt) if shot_noise else 0.0)
