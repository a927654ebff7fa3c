"""Exact FP64 direct sum a_lm = sum_i w_i Y_lm(theta_i) e^{-i m phi_i} -- TEST ONLY.

No spline, no binning: each point's fully normalized associated Legendre
functions are generated by the standard stable recurrence

    Pbar_00 = 1/sqrt(4 pi)
    Pbar_mm = -sqrt((2m+1)/(2m)) sin(theta) Pbar_{m-1,m-1}
    Pbar_{m+1,m} = sqrt(2m+3) cos(theta) Pbar_mm
    Pbar_lm = a_lm (x Pbar_{l-1,m} - b_lm Pbar_{l-2,m}),
      a_lm = sqrt((4l^2-1)/(l^2-m^2)),  b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1))

(Condon-Shortley phase included, Y_lm(theta, phi) = Pbar_lm(cos theta) e^{i m phi}).
This is a different recurrence from the reference's table build
(sht/legendre_py.py:4-39), so agreement between the two is a real check.
The result is the "method-error-free" target: the reference algorithm differs
from it by its cubic-Hermite spline error.

Pinned against scipy.special.sph_harm_y (tests/test_oracle_tables.py);
synth_rings also by the exact Gauss-Legendre quadrature round trip
(tests/test_mocks_host.py).
"""
import numpy as np


def _indx(ell, m, Nl):
    return (m * (2 * Nl - 1 - m)) // 2 + ell


def ylm_values(Nl, theta):
    """Y_lm(theta, 0) for all healpy rows, shape (Nlm, N)."""
    theta = np.asarray(theta, dtype=np.float64)
    x, s = np.cos(theta), np.sin(theta)
    Nlm = Nl * (Nl + 1) // 2
    out = np.empty((Nlm, theta.size))
    pmm = np.full(theta.size, 1.0 / np.sqrt(4 * np.pi))
    for m in range(Nl):
        if m > 0:
            pmm = -np.sqrt((2 * m + 1) / (2.0 * m)) * s * pmm
        out[_indx(m, m, Nl)] = pmm
        if m + 1 < Nl:
            p1 = np.sqrt(2 * m + 3.0) * x * pmm
            out[_indx(m + 1, m, Nl)] = p1
            p2 = pmm
            for ell in range(m + 2, Nl):
                a = np.sqrt((4.0 * ell * ell - 1) / (ell * ell - m * m))
                b = np.sqrt(((ell - 1.0) ** 2 - m * m) / (4.0 * (ell - 1) ** 2 - 1))
                p = a * (x * p1 - b * p2)
                out[_indx(ell, m, Nl)] = p
                p2, p1 = p1, p
    return out


def exact_alm(Nl, theta, phi, w, chunk=4096):
    """sum_i w_i Y*_lm(theta_i, phi_i) in healpy order (complex128)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    Nlm = Nl * (Nl + 1) // 2
    alm = np.zeros(Nlm, dtype=np.complex128)
    for i0 in range(0, theta.size, chunk):
        sl = slice(i0, i0 + chunk)
        Y = ylm_values(Nl, theta[sl])
        for m in range(Nl):
            e = w[sl] * np.exp(-1j * m * phi[sl])
            r = slice(_indx(m, m, Nl), _indx(Nl - 1, m, Nl) + 1)
            alm[r] += Y[r] @ e
    return alm


def exact_alm_multi(Nl, theta, phi, W):
    """exact_alm for several weight vectors W (S, N) sharing the points."""
    return np.stack([exact_alm(Nl, theta, phi, w) for w in np.atleast_2d(W)])


def synth_rings(alm, Nl, x):
    """F[r, m] = sum_{l>=m} a_lm Y_lm(theta_r, 0), cos(theta_r) = x[r]: the
    harmonic synthesis on rings, from ylm_values (shape (len(x), Nl))."""
    Y = ylm_values(Nl, np.arccos(np.clip(np.asarray(x, np.float64), -1, 1)))
    F = np.zeros((len(x), Nl), dtype=np.complex128)
    for m in range(Nl):
        i0 = _indx(m, m, Nl)
        F[:, m] = np.asarray(alm)[i0:i0 + Nl - m] @ Y[i0:i0 + Nl - m]
    return F
