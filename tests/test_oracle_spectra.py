"""CPU checks of the oracle's angular power spectra (alm2cl,
sht/sht.py:95-109, and the cross-spectrum of two alm arrays).  The oracle's
vectorized form is pinned against the reference's own per-l loop form
(written out here as that loop), single-mode closed forms, and for the cross
form against the auto-spectrum through the polarization identity."""
import numpy as np

from oracle import refalg


def rand_alm(rng, Nl):
    Nlm = Nl * (Nl + 1) // 2
    a = rng.normal(size=Nlm) + 1j * rng.normal(size=Nlm)
    ell, m = refalg.getlm(Nl - 1)
    a[m == 0] = a[m == 0].real  # real fields: a_l0 real
    return a


def loop_alm2cl(alm, Nl):
    """sht/sht.py:95-109 as a loop over l."""
    cl = np.zeros(Nl)
    for ell in range(Nl):
        ii = refalg.indx(ell, np.arange(1, ell + 1), Nl)
        cl[ell] = alm[refalg.indx(ell, 0, Nl)].real ** 2 + 2 * np.sum(alm[ii].real ** 2 + alm[ii].imag ** 2)
        cl[ell] /= 2 * ell + 1.0
    return cl


def test_alm2cl_matches_reference_loop_form():
    rng = np.random.default_rng(0)
    for Nl in (1, 2, 17, 64):
        a = rand_alm(rng, Nl)
        np.testing.assert_allclose(refalg.alm2cl(a, Nl), loop_alm2cl(a, Nl), rtol=1e-13)


def test_alm2cl_single_modes():
    Nl = 9
    for ell, m, v in ((0, 0, 3.0), (4, 0, -2.0), (4, 3, 1.5 - 2.0j), (8, 8, 0.5j)):
        a = np.zeros(Nl * (Nl + 1) // 2, complex)
        a[refalg.indx(ell, m, Nl)] = v
        want = np.zeros(Nl)
        want[ell] = (v.real ** 2 if m == 0 else 2 * abs(v) ** 2) / (2 * ell + 1)
        np.testing.assert_allclose(refalg.alm2cl(a, Nl), want, rtol=1e-15, atol=0)


def test_cross_spectrum_identities():
    rng = np.random.default_rng(1)
    Nl = 33
    a, b = rand_alm(rng, Nl), rand_alm(rng, Nl)
    np.testing.assert_allclose(refalg.alm2cl_cross(a, a, Nl), refalg.alm2cl(a, Nl), rtol=1e-14)
    np.testing.assert_allclose(refalg.alm2cl_cross(a, b, Nl), refalg.alm2cl_cross(b, a, Nl), rtol=1e-14)
    polar = (refalg.alm2cl(a + b, Nl) - refalg.alm2cl(a - b, Nl)) / 4
    np.testing.assert_allclose(refalg.alm2cl_cross(a, b, Nl), polar, rtol=1e-10, atol=1e-13)
