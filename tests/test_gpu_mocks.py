"""Harmonic synthesis on rings (dsht_synthesis_rings) and the lognormal mock
generator on the device (directsht_amd.mocks.LogNormalMocks, standing in for
sht/mocks.py without healpy).

* synthesis vs the exact oracle (oracle.exact.synth_rings, pinned to scipy's
  sph_harm_y) at 1e-12;
* map synthesis + Gauss-Legendre quadrature analysis returns the alm (exact
  for nphi >= 2 Nl - 1 and Nl rings);
* the catalogue: Poisson count, seed reproducibility, masking, and the
  clustering of the points (C_l of the point set from the transform) against
  the lognormal theory C_l plus shot noise."""
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

from oracle import exact

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rand_alm(rng, Nl):
    Nlm = Nl * (Nl + 1) // 2
    a = rng.normal(size=Nlm) + 1j * rng.normal(size=Nlm)
    a[:Nl] = a[:Nl].real
    return a


@pytest.mark.parametrize("Nl", [1, 2, 37, 200])
def test_synthesis_vs_oracle(torch_cuda, Nl):
    from directsht_amd import mocks
    rng = np.random.default_rng(Nl)
    a = rand_alm(rng, Nl)
    x = np.concatenate([[1.0, -1.0, 0.0, 0.999999], rng.uniform(-1, 1, 60)])
    F = mocks.synthesize_rings(a, Nl, x).cpu().numpy()
    R = exact.synth_rings(a, Nl, x)
    np.testing.assert_allclose(F, R, rtol=0, atol=1e-12 * np.abs(R).max())


def test_synthesis_quadrature_round_trip(torch_cuda):
    from directsht_amd import mocks
    Nl, nphi = 48, 96
    rng = np.random.default_rng(1)
    a = rand_alm(rng, Nl)
    xr, wr = leggauss(Nl)
    f = mocks.synthesize_map(a, Nl, xr, nphi).cpu().numpy()
    # analysis: a_lm = sum_r w_r Y_lm(x_r) (2pi/nphi) sum_j f_rj e^{-i m phi_j}
    fm = np.fft.fft(f, axis=1)[:, :Nl] * (2 * np.pi / nphi)        # [r][m]
    Y = exact.ylm_values(Nl, np.arccos(xr))                         # [lm][r]
    back = np.zeros_like(a)
    for m in range(Nl):
        i0 = exact._indx(m, m, Nl)
        back[i0:i0 + Nl - m] = Y[i0:i0 + Nl - m] @ (wr * fm[:, m])
    np.testing.assert_allclose(back, a, rtol=0, atol=1e-12 * np.abs(a).max())


def test_gaussian_alm_statistics(torch_cuda):
    torch = torch_cuda
    from directsht_amd import mocks
    Nl = 400
    cl = np.full(Nl, 2.0)
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    a = mocks.gaussian_alm(cl, g).cpu().numpy()
    assert a.shape == (Nl * (Nl + 1) // 2,) and (a[:Nl].imag == 0).all()
    # E|a_lm|^2 = C_l for every m; E(Re a_{l0})^2 = C_l
    assert abs(np.mean(np.abs(a[Nl:]) ** 2) / 2.0 - 1) < 0.01
    assert abs(np.mean(a[:Nl].real ** 2) / 2.0 - 1) < 0.2


def test_lognormal_catalogue(torch_cuda):
    from directsht_amd import mocks
    from directsht_amd.sht import DirectSHT
    Npnt = 1_000_000
    m = mocks.LogNormalMocks(Npnt, nside=128, lmax=128)
    t, p, w = m(seed=7, verbose=False)
    t2, p2, _ = m(seed=7, verbose=False)
    np.testing.assert_array_equal(t, t2)
    np.testing.assert_array_equal(p, p2)
    t3, _, _ = m(seed=8, verbose=False)
    assert t3.size != t.size or not np.array_equal(t3, t)
    m(seed=7, verbose=False)
    t4, _, _ = m.lognormal_catalog(False)   # continues the seeded stream
    assert t4.size != t.size or not np.array_equal(t4, t)
    assert abs(t.size - Npnt) < 6 * np.sqrt(Npnt)
    assert (w == 1).all() and (t > 0).all() and (t < np.pi).all()
    assert (p > 0).all() and (p < 2 * np.pi).all()
    # clustering: C_l of the points (weights 4pi/N) vs lognormal theory + shot noise
    Nl = 100
    sht = DirectSHT(Nl, 2 * Nl)
    alm = sht(t, p, np.full(t.size, 4 * np.pi / t.size))
    cl = sht.alm2cl(alm)
    th = m.get_theory_Cl(lmax_out=Nl - 1)
    ell = np.arange(Nl)
    for lo, hi, tol in [(10, 40, 0.18), (40, 100, 0.10)]:
        band = slice(lo, hi)
        wts = 2 * ell[band] + 1
        ratio = np.sum(wts * cl[band]) / np.sum(wts * th[band])
        assert abs(ratio - 1) < tol, (lo, hi, ratio)
    # a cut catalogue
    mc = mocks.LogNormalMocks(200_000, nside=64, lmax=64, theta_range=(0.5, 1.5),
                              phi_range=(1.0, 3.0))
    tc, pc, _ = mc(seed=1, verbose=False)
    assert tc.size > 0 and (tc > 0.5).all() and (tc < 1.5).all()
    assert (pc > 1.0).all() and (pc < 3.0).all()


def test_synthesis_errors(torch_cuda):
    torch = torch_cuda
    from directsht_amd import _lib
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    buf = torch.zeros(64, dtype=torch.float64, device="cuda")
    for args in [(0, buf.data_ptr(), 1, buf.data_ptr(), buf.data_ptr()),
                 (2049, buf.data_ptr(), 1, buf.data_ptr(), buf.data_ptr()),
                 (4, buf.data_ptr(), -1, buf.data_ptr(), buf.data_ptr()),
                 (4, None, 1, buf.data_ptr(), buf.data_ptr()),
                 (4, buf.data_ptr(), 65535 * 1024 + 1, buf.data_ptr(), buf.data_ptr())]:
        assert lib.dsht_synthesis_rings(*args, st) != 0
        assert b"dsht_synthesis_rings" in lib.dsht_last_error()
    assert lib.dsht_synthesis_rings(4, None, 0, None, None, st) == 0
