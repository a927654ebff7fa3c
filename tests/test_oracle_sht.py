"""Pins for the oracle transform (oracle.refalg.basic_sht / direct_sht).

What pins it: golden alm written by the reference's own C (csht), points on
spline nodes (where the reference algorithm equals the exact direct sum),
coefficients whose Ylm are cubic polynomials in x (reproduced exactly by cubic
Hermite interpolation), linearity, the phi-shift and theta -> pi - theta
symmetries, independence of the polar rotation convention, and the reference's
error behaviour (sht/sht.py:141-149,169-172,205-206).
"""
import numpy as np
import pytest

from oracle import exact, refalg, wigner


def _pts(rng, n, lo=-1.0, hi=1.0):
    t = np.arccos(rng.uniform(lo, hi, n))
    p = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    return t, p, w


@pytest.fixture(scope="module")
def tab64():
    return refalg.make_tables(64, 128, 0.75)


def test_vs_csht_golden(golden_dir):
    d = np.load(f"{golden_dir}/csht_alm_equ.npz")
    T = refalg.make_tables(int(d["Nl_rand"]), int(d["Nx_rand"]), 0.75)
    a = refalg.direct_sht(T, d["theta"], d["phi"], d["wt"])
    ref = d["alm_rand"]
    assert np.max(np.abs(a - ref)) < 1e-13 * np.max(np.abs(ref))
    T = refalg.make_tables(int(d["Nl_det"]), int(d["Nx_det"]), 0.75)
    a = refalg.direct_sht(T, d["theta_det"], d["phi_det"], d["wt_det"])
    ref = d["alm_det"]
    assert np.max(np.abs(a - ref)) < 1e-13 * np.max(np.abs(ref))


def test_vs_csht_golden_band_limit_2100(golden_dir):
    """The oracle against the reference's C at Nl=2100, Nx=40 (table rows and
    equatorial alm); about 25 s on one core."""
    d = np.load(f"{golden_dir}/csht_nl2100.npz")
    xx, Yv, Yd = refalg.make_tables(int(d["Nl"]), int(d["Nx"]), 0.75)
    for mine, ref in ((Yv[d["ridx"]], d["Yv"]), (Yd[d["ridx"]], d["Yd"])):
        scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
        assert np.max(np.abs(mine - ref) / scale) < 1e-12
    a = refalg.direct_sht((xx, Yv, Yd), d["theta"], d["phi"], d["wt"])[d["aidx"]]
    assert np.max(np.abs(a - d["alm"])) / np.max(np.abs(d["alm"])) < 1e-12


def test_nodes_exact(tab64):
    """|cos theta| on a spline node: t = 0 and refalg equals the exact sum."""
    xx = tab64[0]
    rng = np.random.default_rng(21)
    n = 3000
    j = rng.integers(0, 127, n)
    sgn = rng.choice([-1.0, 1.0], n)
    t = np.arccos(sgn * xx[j])
    p = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    a = refalg.direct_sht(tab64, t, p, w)
    e = exact.exact_alm(64, t, p, w)
    assert np.max(np.abs(a - e)) < 1e-13 * np.max(np.abs(e))


def test_cubic_closed_forms(tab64):
    """Even-m Ylm with l <= 3 are cubics in x: Hermite splines are exact."""
    rng = np.random.default_rng(22)
    t, p, w = _pts(rng, 5000, -0.749, 0.749)
    a = refalg.direct_sht(tab64, t, p, w)
    x = np.cos(t)
    closed = {
        (0, 0): np.sum(w) / np.sqrt(4 * np.pi),
        (1, 0): np.sqrt(3 / (4 * np.pi)) * np.sum(w * x),
        (2, 0): np.sqrt(5 / (16 * np.pi)) * np.sum(w * (3 * x ** 2 - 1)),
        (2, 2): np.sqrt(15 / (32 * np.pi)) * np.sum(w * (1 - x ** 2) * np.exp(-2j * p)),
        (3, 0): np.sqrt(7 / (16 * np.pi)) * np.sum(w * (5 * x ** 3 - 3 * x)),
        (3, 2): np.sqrt(105 / (32 * np.pi)) * np.sum(w * x * (1 - x ** 2) * np.exp(-2j * p)),
    }
    for (l, m), v in closed.items():
        assert abs(a[refalg.indx(l, m, 64)] - v) < 1e-12 * abs(a[0]), (l, m)


def test_spline_error_is_small(tab64):
    rng = np.random.default_rng(23)
    t, p, w = _pts(rng, 2000)
    a = refalg.direct_sht(tab64, t, p, w)
    e = exact.exact_alm(64, t, p, w)
    rel = np.max(np.abs(a - e)) / np.max(np.abs(e))
    assert 1e-9 < rel < 1e-4


def test_linearity_and_shift_and_parity(tab64):
    rng = np.random.default_rng(24)
    t, p, w = _pts(rng, 1500)
    w2 = rng.uniform(0.5, 1.5, t.size)
    a1 = refalg.direct_sht(tab64, t, p, w)
    a2 = refalg.direct_sht(tab64, t, p, w2)
    a12 = refalg.direct_sht(tab64, t, p, 2 * w - 3 * w2)
    scale = np.max(np.abs(a1))
    assert np.max(np.abs(a12 - (2 * a1 - 3 * a2))) < 1e-12 * scale
    # phi -> phi + alpha multiplies a_lm by e^{-i m alpha}.  Exact for the
    # equatorial branch only: for polar points a z-shift is an x-rotation in
    # the rotated frame, under which the spline is covariant only up to its
    # interpolation error.
    alpha = 0.7
    ell, m = refalg.getlm(63)
    te, pe, we = _pts(rng, 1500, -0.74, 0.74)
    b1 = refalg.direct_sht(tab64, te, pe, we)
    b_s = refalg.direct_sht(tab64, te, np.mod(pe + alpha, 2 * np.pi), we)
    assert np.max(np.abs(b_s - b1 * np.exp(-1j * m * alpha))) < 1e-12 * np.max(np.abs(b1))
    # theta -> pi - theta multiplies a_lm by (-1)^(l-m)
    a_f = refalg.direct_sht(tab64, np.pi - t, p, w)
    assert np.max(np.abs(a_f - a1 * (-1.0) ** (ell - m))) < 1e-12 * scale


def test_rotation_convention_independence(tab64):
    """R_y(+90), R_y(-90), R_z(a) R_y(90) and R_z(a) R_y(90) R_z(pi) give the
    same alm.  (A pre-rotation R_z(g) with g not in {0, pi} is NOT an exact
    symmetry: it becomes an x-rotation in the rotated frame.  healpy's
    Rotator(rot=(0,90,0), eulertype='ZYZ') has both z-angles zero.)"""
    rng = np.random.default_rng(25)
    t, p, w = _pts(rng, 1500)
    a = refalg.direct_sht(tab64, t, p, w)
    a_m = refalg.direct_sht(tab64, t, p, w, R=wigner.euler_zyz_matrix(0, -np.pi / 2, 0))
    assert np.max(np.abs(a - a_m)) < 1e-13 * np.max(np.abs(a))
    # z-composed: do the polar branch by hand with R_z(a) R_y(90) R_z(g)
    xv = np.abs(np.cos(t))
    pol, equ = xv >= 0.75, xv < 0.75
    for al, ga in ((0.4, 0.0), (-2.1, np.pi)):
        R = wigner.euler_zyz_matrix(al, np.pi / 2, ga)
        tp, pp = refalg.rotate_points(t[pol], p[pol], R)
        b = wigner.rotate_alm_inverse_euler(refalg.basic_sht(tab64, tp, pp, w[pol]), 64, al, 1, ga)
        b += refalg.basic_sht(tab64, t[equ], p[equ], w[equ])
        assert np.max(np.abs(a - b)) < 1e-13 * np.max(np.abs(a))


def test_error_semantics(tab64):
    rng = np.random.default_rng(26)
    t, p, w = _pts(rng, 100, -0.7, 0.7)
    with pytest.raises(AssertionError):
        refalg.basic_sht(tab64, t, p[:-1], w)
    with pytest.raises(AssertionError):
        refalg.basic_sht(tab64, np.array([0.1]), np.array([0.0]), np.array([1.0]))
    with pytest.raises(ValueError):
        refalg.basic_sht(tab64, np.array([]), np.array([]), np.array([]))
    z = refalg.direct_sht(tab64, np.array([]), np.array([]), np.array([]))
    assert z.shape == (64 * 65 // 2,) and np.all(z == 0)


def test_small_xmax_polar_failure_returns_zeros():
    """xmax < sin(pi/4): rotated polar points can stay polar; the reference
    catches the AssertionError and returns zeros (sht/sht.py:141-149)."""
    T = refalg.make_tables(16, 32, 0.6)
    t = np.array([0.9, 1.5, 2.0])  # cos 0.9 > 0.6 and, rotated, |sin 0.9| > 0.6
    p = np.array([0.0, 1.0, 2.0])
    w = np.ones(3)
    a = refalg.direct_sht(T, t, p, w, xmax=0.6)
    assert np.all(a == 0)


def test_exact_c1_fixture_vs_reference_algorithm(golden_dir):
    """The committed exact direct sum of C1 (tests/golden/exact_c1.npz)
    against the reference algorithm on the regenerated points: they differ by
    the spline error alone, 5.8e-7 of max|a| (rms 7.7e-6 of |a|); about 10 s."""
    import sys
    sys.path.insert(0, golden_dir)
    from make_exact_fixtures import c1_points
    d = np.load(f"{golden_dir}/exact_c1.npz")
    t, p, w = c1_points(int(d["seed"]), int(d["n"]))
    a = refalg.direct_sht(refalg.make_tables(256, 512, 0.75), t, p, w)
    ex = d["alm"]
    mx = np.max(np.abs(a - ex)) / np.max(np.abs(ex))
    rms = np.linalg.norm(a - ex) / np.linalg.norm(ex)
    assert 1e-7 < mx < 1e-6 and 1e-6 < rms < 2e-5, (mx, rms)
