"""Pins for the oracle's table build (oracle.refalg.make_tables) and the exact
oracle's per-point Ylm (oracle.exact.ylm_values).

What pins them: scipy's sph_harm_y (an independent library routine), the two
derivative closed forms the reference itself checks (sht/test_sht.py:58-59,
csht/test_alm.py:121-122), the addition theorem, and golden tables written by
the reference's own C code (tests/golden/make_csht_fixtures.py).
"""
import numpy as np
import pytest
from scipy import special

from oracle import exact, refalg


@pytest.fixture(scope="module")
def tab500():
    return refalg.make_tables(500, 1024, 0.75)


def test_grid_matches_reference_construction():
    # sht/sht.py:47-48: arange then in-place multiply by xmax/(Nx-1)
    xx = refalg.x_grid(2048, 0.75)
    assert xx[0] == 0.0 and xx[1] == 0.75 / 2047.0
    assert np.all(xx == np.arange(2048) * (0.75 / 2047.0))


def test_tables_vs_scipy_low_l(tab500):
    """The sht/test_sht.py:21-52 comparison, as an assertion, for l < 45."""
    xx, Yv, Yd = tab500
    th = np.arccos(xx)
    worst = 0.0
    for ell in range(45):
        for m in range(ell + 1):
            ref = special.sph_harm_y(ell, m, th, 0.0).real
            worst = max(worst, np.max(np.abs(Yv[refalg.indx(ell, m, 500)] - ref)))
    assert worst < 1e-13


def test_tables_vs_scipy_high_l():
    Nl, Nx = 200, 256
    xx, Yv, _ = refalg.make_tables(Nl, Nx, 0.75)
    th = np.arccos(xx)
    for ell, m in [(199, 0), (199, 7), (150, 75), (199, 198), (199, 199), (120, 3)]:
        ref = special.sph_harm_y(ell, m, th, 0.0).real
        row = Yv[refalg.indx(ell, m, Nl)]
        assert np.max(np.abs(row - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_derivative_closed_forms(tab500):
    """Reference-held known answers (sht/test_sht.py:57-59)."""
    xx, _, Yd = tab500
    np.testing.assert_allclose(Yd[refalg.indx(2, 0, 500)], 3 * xx * np.sqrt(5 / 4.0 / np.pi),
                               rtol=0, atol=1e-14)
    np.testing.assert_allclose(Yd[refalg.indx(10, 10, 500)],
                               -5.42630291944221461 * xx * (1 - xx ** 2) ** 4, rtol=0, atol=1e-13)


def test_derivative_matches_finite_difference():
    Nl, Nx = 40, 4001
    xx, Yv, Yd = refalg.make_tables(Nl, Nx, 0.75)
    h = xx[1] - xx[0]
    fd = (Yv[:, 2:] - Yv[:, :-2]) / (2 * h)
    scale = np.max(np.abs(Yd), axis=1) + 1.0
    assert np.max(np.max(np.abs(fd - Yd[:, 1:-1]), axis=1) / scale) < 1e-4


def test_addition_theorem(tab500):
    xx, Yv, _ = tab500
    ell, m = refalg.getlm(499)
    wgt = np.where(m == 0, 1.0, 2.0)
    s = np.zeros((500, xx.size))
    np.add.at(s, ell, wgt[:, None] * Yv ** 2)
    expect = (2 * np.arange(500) + 1) / (4 * np.pi)
    assert np.max(np.abs(s / expect[:, None] - 1.0)) < 1e-11


def test_tables_vs_csht_small(golden_dir):
    d = np.load(f"{golden_dir}/csht_tables_small.npz")
    _, Yv, Yd = refalg.make_tables(int(d["Nl"]), int(d["Nx"]), float(d["xmax"]))
    np.testing.assert_allclose(Yv, d["Yv"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(Yd, d["Yd"], rtol=1e-12, atol=1e-13)


def test_tables_vs_csht_rows(golden_dir, tab500):
    d = np.load(f"{golden_dir}/csht_tables_rows.npz")
    _, Yv, Yd = tab500
    idx = np.array([refalg.indx(l, m, 500) for l, m in d["lm"]])
    cols = d["cols"]
    for mine, ref in ((Yv[idx][:, cols], d["Yv"]), (Yd[idx][:, cols], d["Yd"])):
        scale = np.max(np.abs(ref), axis=1, keepdims=True)
        assert np.max(np.abs(mine - ref) / scale) < 1e-11


def test_getlm_roundtrip():
    # sht/shared_utils.py:76-99 against indx (sht/sht.py:58-67)
    for Nl in (1, 2, 7, 64):
        ell, m = refalg.getlm(Nl - 1)
        assert np.all(refalg.indx(ell, m, Nl) == np.arange(Nl * (Nl + 1) // 2))
        assert np.all((m <= ell) & (ell < Nl))


def test_exact_ylm_vs_scipy():
    rng = np.random.default_rng(5)
    th = np.arccos(rng.uniform(-1, 1, 50))
    Nl = 120
    Y = exact.ylm_values(Nl, th)
    for ell, m in [(0, 0), (1, 1), (17, 3), (119, 0), (119, 60), (119, 119), (80, 79)]:
        ref = special.sph_harm_y(ell, m, th, 0.0).real
        assert np.max(np.abs(Y[refalg.indx(ell, m, Nl)] - ref)) < 1e-12


def test_headline_table_rows_vs_csht(golden_dir):
    """Rows of the Nl=1000, Nx=2048 tables at every 16th node against the
    reference's C make_table (tests/golden/csht_c2_equ.npz): the recurrences
    are per column, so the oracle evaluated at those nodes alone gives the
    same values."""
    d = np.load(f"{golden_dir}/csht_c2_equ.npz")
    Nl, Nx = int(d["Nl"]), int(d["Nx"])
    xx = refalg.x_grid(Nx, 0.75)[d["cols"]]
    Yv = refalg.compute_Plm_table(Nl, xx)
    Yd = refalg.compute_der_table(Nl, xx, Yv)
    Yv, Yd = refalg.norm_ext(Yv, Nl), refalg.norm_ext(Yd, Nl)
    for mine, ref in ((Yv[d["ridx"]], d["Yv"]), (Yd[d["ridx"]], d["Yd"])):
        scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
        assert np.max(np.abs(mine - ref) / scale) < 1e-12
