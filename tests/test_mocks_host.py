"""Host-side theory of the lognormal mocks (sht/mocks.py:101-178):
Legendre rows against scipy, the exactness of the Gauss-Legendre transform
pair C_l <-> w(theta), the Gaussian <-> lognormal round trip, the small-field
limit and the shot-noise term.  No GPU."""
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

from directsht_amd import mocks


def test_legendre_rows_vs_scipy():
    sp = pytest.importorskip("scipy.special")
    x = np.linspace(-1, 1, 37)
    P = mocks._legendre_rows(300, x)
    for l in (0, 1, 2, 7, 150, 300):
        np.testing.assert_allclose(P[l], sp.eval_legendre(l, x), rtol=1e-12, atol=1e-14)


def test_transform_pair_is_exact():
    rng = np.random.default_rng(0)
    cl = rng.uniform(0, 1, 81)
    xs, w = leggauss(200)                     # exact for degree < 400
    back = mocks.get_Cl_from_corrfunc(mocks.get_corrfunc_from_Cl(cl, xs), xs, w, 80)
    np.testing.assert_allclose(back, cl, rtol=0, atol=1e-10)
    # w(1) = sum (2l+1) C_l / 4pi (the field variance)
    np.testing.assert_allclose(mocks.get_corrfunc_from_Cl(cl, np.array([1.0]))[0],
                               np.sum((2 * np.arange(81) + 1) * cl) / (4 * np.pi), rtol=1e-13)


def test_lognormal_theory():
    m = mocks.LogNormalMocks(1000, lmax=200, norm=0.01)
    ln = m.get_theory_Cl(shot_noise=False)
    assert ln.shape == (200,)
    # the Gaussian C_l is recovered from the lognormal one (taken far enough
    # past lmax that the lognormal tail is resolved)
    ln_full = m.get_theory_Cl(lmax_out=599, shot_noise=False)
    np.testing.assert_allclose(ln_full[:200], ln, rtol=1e-12)
    g = mocks.get_gauss_cl_from_ln_cl(ln_full)
    np.testing.assert_allclose(g[:200], m.clg, rtol=1e-7)
    assert np.abs(g[200:]).max() < 1e-9 * m.clg.max()
    # small fields: exp(w) - 1 ~ w, so C_l(LN) ~ C_l(G)
    small = mocks.LogNormalMocks(1000, lmax=60, norm=1e-7)
    np.testing.assert_allclose(small.get_theory_Cl(shot_noise=False), small.clg, rtol=1e-5)
    # shot noise 4pi / Npnt
    np.testing.assert_allclose(m.get_theory_Cl() - ln, 4 * np.pi / 1000, rtol=1e-12)
    np.testing.assert_allclose(m.get_theory_Cl(lmax_out=50, shot_noise=False), ln[:51], rtol=1e-12)
    with pytest.raises(AssertionError):
        mocks.LogNormalMocks(10, clg=np.ones(5), cl_ln=np.ones(5))
    c = mocks.LogNormalMocks(10, cl_ln=ln_full)
    np.testing.assert_allclose(c.clg[:200], m.clg, rtol=1e-7)


def test_mask():
    m = mocks.LogNormalMocks(10, theta_range=(0.5, 1.0), phi_range=(1.0, 2.0))
    t = np.array([0.4, 0.6, 0.6, 0.9, 1.0])
    p = np.array([1.5, 1.5, 2.5, 1.1, 1.5])
    np.testing.assert_array_equal(m.make_mask(t, p), [1, 3])


def test_synth_rings_oracle_quadrature_round_trip():
    """oracle.exact.synth_rings: the map it synthesizes on Nl Gauss-Legendre
    rings and 2Nl - 1 longitudes returns its alm under exact quadrature."""
    from oracle import exact
    Nl = 24
    nphi = 2 * Nl - 1
    rng = np.random.default_rng(2)
    Nlm = Nl * (Nl + 1) // 2
    a = rng.normal(size=Nlm) + 1j * rng.normal(size=Nlm)
    a[:Nl] = a[:Nl].real
    xr, wr = leggauss(Nl)
    F = exact.synth_rings(a, Nl, xr)
    phi = 2 * np.pi * np.arange(nphi) / nphi
    m = np.arange(Nl)
    f = (F[:, :1].real + 2 * (F[:, 1:, None] * np.exp(1j * m[1:, None] * phi)).real.sum(axis=1))
    fm = np.fft.fft(f, axis=1)[:, :Nl] * (2 * np.pi / nphi)
    Y = exact.ylm_values(Nl, np.arccos(xr))
    back = np.zeros_like(a)
    for mm in range(Nl):
        i0 = exact._indx(mm, mm, Nl)
        back[i0:i0 + Nl - mm] = Y[i0:i0 + Nl - mm] @ (wr * fm[:, mm])
    np.testing.assert_allclose(back, a, rtol=0, atol=1e-12 * np.abs(a).max())
