"""FP64 numpy restatement of the reference direct-SHT algorithm ("refalg").

TEST INFRASTRUCTURE ONLY -- never imported by the product path.

Every function cites the reference file:line it follows (paths relative to the
reference checkout).  The arithmetic of the table recurrences is written in the
same order as the reference's numba loops so the tables agree to the last bit
or two; the binning, v's and contraction follow the reference's per-bin form
(not the node-basis fold the GPU kernels use), so the GPU path is checked
against an algebraically equivalent but independently ordered computation.

Parity pins (tests/test_oracle_tables.py, tests/test_oracle_sht.py):
  * tables: scipy ``sph_legendre_p`` for l < 200, the reference-held closed
    forms dY_20/dx and dY_{10,10}/dx (sht/test_sht.py:58-59), the addition
    theorem sum_m |Y_lm|^2 (1+[m>0]) = (2l+1)/4pi, and csht golden fixtures;
  * basic_sht / direct_sht: nodes-exact points against ``oracle.exact``,
    closed forms (a_00, a_10, a_20, a_22) that cubic Hermite reproduces
    exactly, linearity, phi-shift and theta -> pi - theta invariants, and
    csht golden fixtures for the equatorial path;
  * the polar branch: rotation of the exact-oracle alm (oracle.wigner), and
    independence of the result from the rotation convention;
  * alm2cl: the reference's own per-l loop form (sht/sht.py:95-109), closed
    forms for single modes, and invariance under rotation (oracle.wigner);
    alm2cl_cross: cross(a, a) = alm2cl(a), symmetry, and the polarization
    identity cross(a, b) = [alm2cl(a + b) - alm2cl(a - b)] / 4.
"""
import numpy as np

from . import wigner

__all__ = [
    "indx", "getlm", "x_grid", "compute_Plm_table", "compute_der_table",
    "norm_ext", "make_tables", "basic_sht", "direct_sht", "rotate_points",
    "ROT_Y_PLUS_90", "alm2cl", "alm2cl_cross",
]

# R_y(+90 deg), active: n' = R n maps (x, y, z) -> (z, y, -x).  Every rotation
# R_z(a) R_y(+-90 deg) R_z(g) with g in {0, pi} gives the same final alm to
# rounding: in the rotated frame these differ by a z-rotation (exact phases),
# R_y(180) (theta -> pi - theta, phi -> pi - phi) or R_x(180), all exact
# symmetries of the Hermite approximant.  healpy's
# Rotator(rot=(0,90,0), eulertype='ZYZ') (sht/sht.py:134,138) has both z-angles
# zero, so this restatement does not need healpy's sign conventions.
ROT_Y_PLUS_90 = np.array([[0.0, 0.0, 1.0],
                          [0.0, 1.0, 0.0],
                          [-1.0, 0.0, 0.0]])


def indx(ell, m, Nl):
    """healpy index of (ell, m) -- sht/sht.py:58-67, csht/sht_helper.c:8-13."""
    return (m * (2 * Nl - 1 - m)) // 2 + ell


def getlm(lmax, szalm=None):
    """(ell, m) of every healpy index -- sht/shared_utils.py:76-99."""
    if szalm is None:
        szalm = (lmax + 1) * (lmax + 2) // 2
    i = np.arange(szalm)
    m = np.ceil(((2 * lmax + 1) - np.sqrt((2 * lmax + 1) ** 2 - 8 * (i - lmax))) / 2).astype(int)
    ell = i - m * (2 * lmax + 1 - m) // 2
    return ell, m


def x_grid(Nx, xmax):
    """Spline nodes x_j = j * (xmax/(Nx-1)) -- sht/sht.py:47-48."""
    xx = np.arange(Nx, dtype="float64")
    xx *= xmax / float(Nx - 1)
    return xx


def compute_Plm_table(Nl, xx):
    """Table of sqrt(4pi/(2l+1)) Y_lm(x, 0) in healpy row order.

    Follows sht/legendre_py.py:4-39 (m=0 seed, P_mm, P_{m+1,m}) and the
    ext_slow_recurrence loop sht/legendre_py.py:84-101.  The l >= m+2 loop is
    vectorized over m (rows of one m only depend on rows of the same m), with
    the same per-element operation order as the numba loop.
    """
    Nx = len(xx)
    sx = np.sqrt(1 - xx ** 2)
    Nlm = (Nl * (Nl + 1)) // 2
    Plm = np.zeros((Nlm, Nx), dtype="float64")
    Plm[indx(0, 0, Nl), :] = np.ones_like(xx)
    if Nl > 1:
        Plm[indx(1, 0, Nl), :] = xx.copy()
    # m = 0, l >= 2 (legendre_py.py:21-24); overwritten below exactly as the
    # reference's ext_slow_recurrence (which starts at m = 0) overwrites it.
    for ell in range(2, Nl):
        i0, i1, i2 = indx(ell, 0, Nl), indx(ell - 1, 0, Nl), indx(ell - 2, 0, Nl)
        Plm[i0, :] = (2 * ell - 1) * xx * Plm[i1, :] - (ell - 1) * Plm[i2, :]
        Plm[i0, :] /= float(ell)
    # P_mm (legendre_py.py:28-30)
    for m in range(1, Nl):
        Plm[indx(m, m, Nl), :] = -np.sqrt(1.0 - 1.0 / (2 * m)) * sx * Plm[indx(m - 1, m - 1, Nl), :]
    # P_{m+1,m} (legendre_py.py:32-34)
    for m in range(1, Nl - 1):
        Plm[indx(m + 1, m, Nl), :] = np.sqrt(2 * m + 1.0) * xx * Plm[indx(m, m, Nl), :]
    # l >= m+2 (legendre_py.py:91-100), vectorized over m at fixed k = l - m
    for k in range(2, Nl):
        m = np.arange(0, Nl - k)
        ell = m + k
        i0, i1, i2 = indx(ell, m, Nl), indx(ell - 1, m, Nl), indx(ell - 2, m, Nl)
        fact1 = np.sqrt((ell - m) * 1.0 / (ell + m))
        fact2 = np.sqrt((ell - m - 1.0) / (ell + m - 1.0))
        a = (2 * ell - 1)[:, None] * xx[None, :] * Plm[i1, :]
        b = (ell + m - 1)[:, None] * Plm[i2, :] * fact2[:, None]
        Plm[i0, :] = a - b
        Plm[i0, :] *= (fact1 / (ell - m))[:, None]
    return Plm


def compute_der_table(Nl, xx, Yv):
    """d/dx of the (unnormalized) table -- sht/legendre_py.py:41-61,104-118.

    The l = m rows read row indx(m-1, m) (= indx(Nl-1, m-1)) times a zero
    factor, as the reference does.
    """
    Nlm = (Nl * (Nl + 1)) // 2
    Yd = np.zeros((Nlm, xx.size), dtype="float64")
    if Nl > 1:
        Yd[indx(1, 0, Nl), :] = np.ones_like(xx)
    for ell in range(2, Nl):
        i0, i1 = indx(ell, 0, Nl), indx(ell - 1, 0, Nl)
        Yd[i0, :] = ell / (1 - xx ** 2) * (Yv[i1, :] - xx * Yv[i0, :])
    omx2 = 1.0 - xx ** 2
    for m in range(1, Nl):
        ell = np.arange(m, Nl)
        i0, i1 = indx(ell, m, Nl), indx(ell - 1, m, Nl)
        fact = np.sqrt((ell - m).astype(float) / (ell + m))
        Yd[i0, :] = ((ell + m) * fact)[:, None] * Yv[i1, :] - (ell[:, None] * xx[None, :]) * Yv[i0, :]
        Yd[i0, :] /= omx2
    return Yd


def norm_ext(Y, Nl):
    """Multiply row (l, m) by sqrt((2l+1)/4pi) -- sht/legendre_py.py:67-83."""
    ell, _ = getlm(Nl - 1)
    fact = np.sqrt((2 * ell + 1) / 4.0 / np.pi)
    Y *= fact[:, None]
    return Y


def make_tables(Nl, Nx, xmax=0.75):
    """(x, Yv, Yd) exactly as DirectSHT.__init__ builds them -- sht/sht.py:44-56."""
    xx = x_grid(Nx, xmax)
    Yv = compute_Plm_table(Nl, xx)
    Yd = compute_der_table(Nl, xx, Yv)
    Yv, Yd = norm_ext(Yv, Nl), norm_ext(Yd, Nl)
    return xx, Yv, Yd


def _hemisphere_alm(Yv, Yd, xx, Nl, x, phi, w, m_chunk=64):
    """alm of one hemisphere (x >= 0 already folded), sht/sht.py:209-277.

    Binning: np.digitize (sht/sht.py:212), t (sht/sht.py:221), Hermite inputs
    (sht/sht.py:239-241); v's as sht/shared_interp_funcs.py:88-105 /
    sht/interp_funcs_py.py:16-44 (sums over the points of each bin; this
    restatement sums each bin's points with np.add.reduceat instead of the
    zero-padded 2-D layout, which only changes the summation order); the
    contraction is sht/interp_funcs_py.py:46-60 with vs = vs_r - 1j vs_i
    (sht/sht.py:271) and dx folded into the derivative table (sht/sht.py:276).
    """
    Nlm = (Nl * (Nl + 1)) // 2
    dx = xx[1] - xx[0]
    spline_idx = np.digitize(x, xx) - 1
    order = np.argsort(spline_idx, kind="stable")
    spline_idx = spline_idx[order]
    xs, ws, ps = x[order], w[order], phi[order]
    t = (xs - xx[spline_idx]) / dx
    inputs = np.vstack([ws * (2 * t + 1) * (1 - t) ** 2, ws * t * (1 - t) ** 2,
                        ws * t ** 2 * (3 - 2 * t), ws * t ** 2 * (t - 1)])
    occupied, starts = np.unique(spline_idx, return_index=True)
    nb = len(occupied)
    vs_r = np.zeros((Nl, 4, nb))
    vs_i = np.zeros((Nl, 4, nb))
    for m0 in range(0, Nl, m_chunk):
        ms = np.arange(m0, min(Nl, m0 + m_chunk), dtype="float64")
        arg = np.outer(ms, ps)
        c, s = np.cos(arg), np.sin(arg)
        for k in range(4):
            vs_r[m0:m0 + len(ms), k, :] = np.add.reduceat(c * inputs[k], starts, axis=1)
            vs_i[m0:m0 + len(ms), k, :] = np.add.reduceat(s * inputs[k], starts, axis=1)
    vs = vs_r - 1j * vs_i
    alm = np.zeros(Nlm, dtype="complex128")
    Y0, Y1 = Yv[:, occupied], Yv[:, occupied + 1]
    D0, D1 = dx * Yd[:, occupied], dx * Yd[:, occupied + 1]
    for m in range(Nl):
        r = slice(indx(m, m, Nl), indx(Nl - 1, m, Nl) + 1)
        alm[r] = Y0[r] @ vs[m, 0] + D0[r] @ vs[m, 1] + Y1[r] @ vs[m, 2] + D1[r] @ vs[m, 3]
    return alm


def basic_sht(tables, theta, phi, wt, xmax=0.75, reg_factor=1.0):
    """DirectSHT.basic_sht -- sht/sht.py:153-283 (FP64).

    ``tables`` = (x, Yv, Yd) from make_tables.  Raises AssertionError on a
    length mismatch or theta outside (arccos(xmax), arccos(-xmax))
    (sht/sht.py:169-172) and ValueError on empty input (sht/sht.py:205-206).
    """
    xx, Yv, Yd = tables
    Nlm = Yv.shape[0]
    Nl = int(round((np.sqrt(8 * Nlm + 1) - 1) / 2))
    theta, phi, wt = (np.asarray(a, dtype="float64") for a in (theta, phi, wt))
    assert len(theta) == len(phi) and len(phi) == len(wt), "theta,phi,wt must be the same length."
    assert np.all((theta >= 0) & (theta > np.arccos(xmax)) & (theta < np.arccos(-xmax))), \
        "theta must be in [ACos[xmax],ACos[-xmax])."
    wt = wt * reg_factor
    ell, m = getlm(Nl - 1, Nlm)
    parity = (-1.0) ** (ell - m)
    x_full = np.cos(theta)
    pos, neg = np.where(x_full >= 0)[0], np.where(x_full < 0)[0]
    if len(pos) == 0 and len(neg) == 0:
        raise ValueError("The theta array seems to be empty!")
    alm = np.zeros(Nlm, dtype="complex128")
    if len(pos) > 0:
        alm += _hemisphere_alm(Yv, Yd, xx, Nl, x_full[pos], phi[pos], wt[pos])
    if len(neg) > 0:
        alm += parity * _hemisphere_alm(Yv, Yd, xx, Nl, np.abs(x_full[neg]), phi[neg], wt[neg])
    return alm / reg_factor


def rotate_points(theta, phi, R=ROT_Y_PLUS_90):
    """Active rotation n' = R n of points given as (theta, phi).

    Stands in for healpy ``Rotator(rot=(0,90,0), eulertype='ZYZ', deg=True)``
    (sht/sht.py:134-135); angles are recovered the way healpy's vec2ang does
    (theta = arctan2(rho, z), phi = arctan2(y, x) wrapped to [0, 2pi)).
    """
    st = np.sin(theta)
    v = np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)])
    vr = R @ v
    tr = np.arctan2(np.sqrt(vr[0] ** 2 + vr[1] ** 2), vr[2])
    pr = np.arctan2(vr[1], vr[0])
    pr = np.where(pr < 0.0, pr + 2 * np.pi, pr)
    return tr, pr


def direct_sht(tables, t, p, w, xmax=0.75, R=ROT_Y_PLUS_90, verbose=False):
    """DirectSHT.__call__ -- sht/sht.py:112-150.

    Polar points (|cos t| >= xmax) are rotated by R, transformed with
    basic_sht, and the result is rotated back with the Wigner-d rotation of
    ``oracle.wigner`` (healpy's ``rotate_alm`` in the reference).  As in the
    reference, an AssertionError in the polar basic_sht (possible when
    xmax < sin(pi/4)) is caught and the equatorial part is then skipped, so
    the result is zero (sht/sht.py:141-149).
    """
    xx, Yv, Yd = tables
    Nlm = Yv.shape[0]
    Nl = int(round((np.sqrt(8 * Nlm + 1) - 1) / 2))
    t, p, w = (np.asarray(a, dtype="float64") for a in (t, p, w))
    alms = np.zeros(Nlm, dtype="complex128")
    xval = np.abs(np.cos(t))
    pol = np.nonzero(xval >= xmax)[0]
    equ = np.nonzero(xval < xmax)[0]
    ok = True
    if len(pol) > 0:
        ok = False
        tp, pp = rotate_points(t[pol], p[pol], R)
        try:
            rlms = basic_sht(tables, tp, pp, w[pol], xmax=xmax)
            alms += wigner.rotate_alm_inverse(rlms, Nl, R)
            ok = True
        except AssertionError as ae:
            if verbose:
                print(f"AssertionError in basic_sht: {ae}")
    if ok and len(equ) > 0:
        alms += basic_sht(tables, t[equ], p[equ], w[equ], xmax=xmax)
    return alms


def alm2cl(alm, Nl):
    """Pseudo-spectrum 1/(2l+1) sum_m |a_lm|^2 -- sht/sht.py:95-109."""
    ell, m = getlm(Nl - 1)
    p = np.where(m == 0, alm.real ** 2, 2 * (alm.real ** 2 + alm.imag ** 2))
    cl = np.bincount(ell, weights=p, minlength=Nl)
    return cl / (2 * np.arange(Nl) + 1.0)


def alm2cl_cross(a, b, Nl):
    """Cross-spectrum 1/(2l+1) [Re a_l0 Re b_l0 + 2 sum_{m>0} Re(a_lm conj(b_lm))],
    healpy's alm2cl(alm1, alm2) as the reference's callers use it
    (notebooks/analyzing_mocks.ipynb) with alm2cl's m = 0 convention
    (sht/sht.py:95-109); alm2cl_cross(a, a) = alm2cl(a)."""
    ell, m = getlm(Nl - 1)
    p = np.where(m == 0, a.real * b.real, 2 * (a.real * b.real + a.imag * b.imag))
    cl = np.bincount(ell, weights=p, minlength=Nl)
    return cl / (2 * np.arange(Nl) + 1.0)
