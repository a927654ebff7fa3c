"""Pins for oracle.wigner (the stand-in for healpy's rotate_alm, sht/sht.py:138-139).

What pins it: the explicit factorial-sum formula for Wigner d evaluated in
40-digit arithmetic (mpmath), orthogonality at high l, and the mathematical
fact that rotating the exact alm of rotated points must give the exact alm of
the original points.
"""
import mpmath as mp
import numpy as np
import pytest

from oracle import exact, refalg, wigner


def _d_explicit(j, mp_, m, beta):
    mp.mp.dps = 40
    c, s = mp.cos(beta / 2), mp.sin(beta / 2)
    pref = mp.sqrt(mp.factorial(j + mp_) * mp.factorial(j - mp_) * mp.factorial(j + m)
                   * mp.factorial(j - m))
    tot = mp.mpf(0)
    for k in range(max(0, m - mp_), min(j + m, j - mp_) + 1):
        den = (mp.factorial(j + m - k) * mp.factorial(k) * mp.factorial(mp_ - m + k)
               * mp.factorial(j - mp_ - k))
        tot += (-1) ** (mp_ - m + k) / den * c ** (2 * j + m - mp_ - 2 * k) * s ** (mp_ - m + 2 * k)
    return float(pref * tot)


@pytest.mark.parametrize("beta_sign", [1, -1])
def test_d_vs_explicit_formula(beta_sign):
    for l in (0, 1, 2, 5, 12):
        D = wigner.d_beta(l, beta_sign)
        for mp_ in range(-l, l + 1):
            for m in range(0, l + 1):
                ref = _d_explicit(l, mp_, m, beta_sign * mp.pi / 2)
                assert abs(D[mp_ + l, m] - ref) < 1e-14, (l, mp_, m)


def test_d1_sign_convention():
    D = wigner.d_halfpi(1)
    assert abs(D[2, 0] + np.sqrt(0.5)) < 1e-15  # d^1_{10}(pi/2) = -1/sqrt(2)


@pytest.mark.parametrize("l", [64, 700, 1100])
def test_orthogonality(l):
    D = wigner.d_halfpi(l)
    assert np.max(np.abs(D.T @ D - np.eye(l + 1))) < 1e-12


@pytest.mark.parametrize("alpha,beta_sign,gamma", [(0.0, 1, 0.0), (0.0, -1, 0.0),
                                                   (0.3, -1, 1.1), (2.0, 1, -0.7)])
def test_rotation_of_exact_alm(alpha, beta_sign, gamma):
    rng = np.random.default_rng(11)
    N, Nl = 400, 48
    t = np.arccos(rng.uniform(-1, 1, N))
    p = rng.uniform(0, 2 * np.pi, N)
    w = rng.uniform(0.5, 1.5, N)
    a = exact.exact_alm(Nl, t, p, w)
    R = wigner.euler_zyz_matrix(alpha, beta_sign * np.pi / 2, gamma)
    tr, pr = refalg.rotate_points(t, p, R)
    back = wigner.rotate_alm_inverse_euler(exact.exact_alm(Nl, tr, pr, w), Nl, alpha, beta_sign, gamma)
    assert np.max(np.abs(back - a)) < 1e-13 * np.max(np.abs(a))


def test_rotation_preserves_cl():
    rng = np.random.default_rng(12)
    Nl = 40
    a = rng.normal(size=Nl * (Nl + 1) // 2) + 1j * rng.normal(size=Nl * (Nl + 1) // 2)
    ell, m = refalg.getlm(Nl - 1)
    a[m == 0] = a[m == 0].real
    b = wigner.rotate_alm_inverse(a, Nl, refalg.ROT_Y_PLUS_90)
    np.testing.assert_allclose(refalg.alm2cl(b, Nl), refalg.alm2cl(a, Nl), rtol=1e-12)
