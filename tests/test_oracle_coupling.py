"""Pins of the Wigner-3j and mode-coupling oracle (oracle/coupling.py) on the
CPU: the reference's compiled C table, sympy's exact symbols at the points the
reference's own test checks (sht/test_3j.py:20-40), the Racah closed form, the
orthogonality sum, and the mode-coupling identities of
sht/mask_deconvolution.py (full-sky debug window -> identity)."""
from fractions import Fraction
from math import factorial

import numpy as np
import pytest

from oracle import coupling, cref


def racah_sq(l1, l2, l3):
    """(l1 l2 l3; 0 0 0)^2 as an exact rational (zero off the triangle or for
    odd l1 + l2 + l3)."""
    J = l1 + l2 + l3
    if J % 2 or l3 < abs(l1 - l2) or l3 > l1 + l2:
        return Fraction(0)
    g = J // 2
    r = Fraction(factorial(g) ** 2, factorial(2 * g + 1))
    for l in (l1, l2, l3):
        r *= Fraction(factorial(2 * g - 2 * l), factorial(g - l) ** 2)
    return r


def test_index_is_the_packed_order():
    N = 9
    seen = []
    for j1 in range(N):
        for j2 in range(j1 + 1):
            for j3 in range(j2 + 1):
                seen.append(coupling.tj_index(j1, j2, j3))
                assert coupling.tj_index(j3, j1, j2) == seen[-1]
    assert seen == list(range(N * (N + 1) * (N + 2) // 6))


@pytest.mark.skipif(not cref.threej_available(), reason="oracle/_ref not built")
def test_table_equals_reference_c():
    N = 60
    ref = cref.threej000_table(N)
    np.testing.assert_array_equal(coupling.threej000_table(N, prefill=False), ref)
    pre = coupling.threej000_table(N, prefill=True)
    np.testing.assert_allclose(pre, ref, rtol=2e-15, atol=0)
    # the scalar recursion agrees with the table, both variants
    rng = np.random.default_rng(0)
    for j1, j2, j3 in rng.integers(0, N, size=(200, 3)):
        i = coupling.tj_index(j1, j2, j3)
        assert coupling.threej000(j1, j2, j3) == ref[i]
        assert coupling.threej000(j1, j2, j3, prefill=True) == pre[i]


def test_table_against_racah_and_sympy():
    sympy_w = pytest.importorskip("sympy.physics.wigner")
    N = 40
    tab = coupling.threej000_table(N)
    for j1 in range(N):
        for j2 in range(j1 + 1):
            for j3 in range(j2 + 1):
                v = tab[coupling.tj_index(j1, j2, j3)]
                want = racah_sq(j1, j2, j3)
                assert abs(v * v - float(want)) <= 1e-14 * float(want) + 1e-300
                if want and (j1 + j2 + j3) // 2 % 2:
                    assert v < 0
                elif want:
                    assert v > 0
    for j1 in range(12):
        for j2 in range(j1 + 1):
            for j3 in range(j2 + 1):
                ex = float(sympy_w.wigner_3j(j1, j2, j3, 0, 0, 0))
                assert abs(tab[coupling.tj_index(j1, j2, j3)] - ex) <= 4e-15 * abs(ex) + 1e-300


def test_reference_test_points():
    """sht/test_3j.py:20-40 compares Wigner3j(1000) with sympy at
    (999, 998, 1), (998, 997, 5) and (l, l+4, l+8), l = 2, 8, 20."""
    sympy_w = pytest.importorskip("sympy.physics.wigner")
    pts = [(999, 998, 1), (998, 997, 5), (2, 6, 10), (8, 12, 16), (20, 24, 28)]
    for l1, l2, l3 in pts:
        ex = float(sympy_w.wigner_3j(l1, l2, l3, 0, 0, 0).n(32))
        got = coupling.threej000(l1, l2, l3, prefill=True)
        assert abs(got - ex) <= 1e-13 * abs(ex), (l1, l2, l3, got, ex)


def test_orthogonality():
    N = 61
    T = coupling.dense_threej000(N)
    l3 = np.arange(N)
    for l1 in range(0, 30):
        for l2 in range(0, 30):
            s = np.sum((2 * l3 + 1) * T[l1, l2, :] ** 2)
            np.testing.assert_allclose(s, 1.0, rtol=1e-14)


def test_mode_coupling_identities():
    Nl = 24
    rng = np.random.default_rng(5)
    # full-sky debug window (sht/mask_deconvolution.py:162-167): identity
    np.testing.assert_allclose(coupling.mode_coupling(Nl, np.zeros(3), debug=True), np.eye(Nl),
                               atol=1e-14)
    # constant window over every l3 reached: (2 l2 + 1) c / 4pi in every row
    c = 0.37
    M = coupling.mode_coupling(Nl, np.full(2 * Nl - 1, c))
    want = np.broadcast_to((2 * np.arange(Nl) + 1) * c / (4 * np.pi), (Nl, Nl))
    np.testing.assert_allclose(M, want, rtol=1e-13)
    # linearity, reciprocity, and the zero padding of a short window
    W1, W2 = rng.uniform(0, 1, 2 * Nl - 1), rng.uniform(-1, 1, 2 * Nl - 1)
    M1, M2 = coupling.mode_coupling(Nl, W1), coupling.mode_coupling(Nl, W2)
    M12 = coupling.mode_coupling(Nl, 2.0 * W1 - 3.0 * W2)
    np.testing.assert_allclose(M12, 2 * M1 - 3 * M2, atol=1e-13 * np.abs(M12).max())
    ell = np.arange(Nl)
    np.testing.assert_allclose((2 * ell + 1)[:, None] * M1, ((2 * ell + 1)[:, None] * M1).T,
                               rtol=1e-13)
    short = W1.copy()
    short[10:] = 0
    np.testing.assert_array_equal(coupling.mode_coupling(Nl, W1[:10]),
                                  coupling.mode_coupling(Nl, short))
    # one entry against exact rationals
    l1, l2 = 7, 11
    ex = sum((2 * l3 + 1) * Fraction(float(W1[l3])) * racah_sq(l1, l2, l3)
             for l3 in range(2 * Nl - 1))
    np.testing.assert_allclose(M1[l1, l2], float(ex) * (2 * l2 + 1) / (4 * np.pi), rtol=1e-14)
