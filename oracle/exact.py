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
    """Y_lm(theta, 0) for all healpy rows, shape (Nlm, N).  The recurrence in
    l runs for all orders m at once (rows with the same l - m)."""
    theta = np.asarray(theta, dtype=np.float64)
    x, s = np.cos(theta), np.sin(theta)
    Nlm = Nl * (Nl + 1) // 2
    out = np.empty((Nlm, theta.size))
    ms = np.arange(Nl)
    diag = _indx(ms, ms, Nl)
    pmm = np.full(theta.size, 1.0 / np.sqrt(4 * np.pi))
    out[diag[0]] = pmm
    for m in range(1, Nl):
        pmm = -np.sqrt((2 * m + 1) / (2.0 * m)) * s * pmm
        out[diag[m]] = pmm
    if Nl > 1:
        m = ms[:-1]
        out[_indx(m + 1, m, Nl)] = np.sqrt(2 * m + 3.0)[:, None] * x[None, :] * out[diag[:-1]]
    for k in range(2, Nl):
        m = ms[:Nl - k]
        ell = (m + k).astype(np.float64)
        a = np.sqrt((4.0 * ell * ell - 1) / (ell * ell - m * m))
        b = np.sqrt(((ell - 1.0) ** 2 - m * m) / (4.0 * (ell - 1) ** 2 - 1))
        out[_indx(m + k, m, Nl)] = a[:, None] * (x[None, :] * out[_indx(m + k - 1, m, Nl)]
                                                 - b[:, None] * out[_indx(m + k - 2, m, Nl)])
    return out


def exact_alm(Nl, theta, phi, w, chunk=None):
    """sum_i w_i Y*_lm(theta_i, phi_i) in healpy order (complex128); w may be
    (S, N) weight rows sharing the points, giving (S, Nlm).  Points go
    through the Ylm recurrence `chunk` at a time (default: about 1 GB of
    Ylm values per chunk)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    one = w.ndim == 1
    W = np.atleast_2d(w)
    Nlm = Nl * (Nl + 1) // 2
    if chunk is None:
        chunk = int(max(16, min(4096, (1 << 27) // max(Nlm, 1))))
    alm = np.zeros((W.shape[0], Nlm), dtype=np.complex128)
    for i0 in range(0, theta.size, chunk):
        sl = slice(i0, i0 + chunk)
        Y = ylm_values(Nl, theta[sl])
        for m in range(Nl):
            e = W[:, sl] * np.exp(-1j * m * phi[sl])[None, :]
            r = slice(_indx(m, m, Nl), _indx(Nl - 1, m, Nl) + 1)
            alm[:, r] += e @ Y[r].T
    return alm[0] if one else alm


def exact_alm_multi(Nl, theta, phi, W):
    """exact_alm for several weight vectors W (S, N) sharing the points."""
    return exact_alm(Nl, theta, phi, np.atleast_2d(W))


def synth_rings(alm, Nl, x):
    """F[r, m] = sum_{l>=m} a_lm Y_lm(theta_r, 0), cos(theta_r) = x[r]: the
    harmonic synthesis on rings, from ylm_values (shape (len(x), Nl))."""
    Y = ylm_values(Nl, np.arccos(np.clip(np.asarray(x, np.float64), -1, 1)))
    F = np.zeros((len(x), Nl), dtype=np.complex128)
    for m in range(Nl):
        i0 = _indx(m, m, Nl)
        F[:, m] = np.asarray(alm)[i0:i0 + Nl - m] @ Y[i0:i0 + Nl - m]
    return F
