"""Wigner small-d at beta = +-pi/2 and the alm rotation -- TEST INFRASTRUCTURE.

Replaces healpy ``Rotator(...).rotate_alm`` on the reference's polar branch
(sht/sht.py:138-139; healpy is a third-party dependency, unpinned in the
reference's setup.py:22, and is not installed here).

Convention (standard / Edmonds): d^l_{m'm}(b) = <l m'| exp(-i b J_y) |l m>,
so d^1_{10}(b) = -sin(b)/sqrt(2).  For points rotated actively by R,
n' = R n, with R = R_z(alpha) R_y(beta) R_z(gamma),

    a_lm = sum_{m'=-l..l} exp(i m' alpha) d^l_{m'm}(beta) exp(i m gamma) a'_{lm'},

with a'_{l,-m'} = (-1)^{m'} conj(a'_{lm'}) for real weights.

Algorithm (Trapani & Navaza, Acta Cryst. A62 (2006) 262, specialised here to
a downward three-term recursion in the first index at fixed second index m):

    d_{l,m}(pi/2)      = (-1)^{l-m} sqrt(C(2l, l+m)) / 2^l
    d_{m'-1,m}(pi/2)   = [2m d_{m',m} - sqrt((l-m')(l+m'+1)) d_{m'+1,m}]
                         / sqrt((l+m')(l-m'+1))
    d_{-m',m}(pi/2)    = (-1)^{l+m} d_{m',m}(pi/2)

The recursion runs from m' = l (classically forbidden, tiny values) down to
m' = 0, i.e. always in the direction in which the wanted solution grows, and
carries a power-of-two exponent so that starting values below the double
range (2^-l for l > 1022) do not flush the column to zero.

Parity pins (tests/test_oracle_wigner.py): the explicit factorial sum
evaluated in 40-digit mpmath for l <= 12 (all m', m, both signs of beta),
orthogonality sum_m' d_{m'm} d_{m'k} = delta_mk at l up to 1100, and the
rotation of exact-oracle alm against the exact alm of the rotated points.
"""
import numpy as np

_SCALE_HI = 2.0 ** 400
_SCALE_LO = 2.0 ** -400


def _top_row(l):
    """(mantissa, exponent) of d^l_{l,m}(pi/2) for m = 0..l."""
    mant = np.empty(l + 1)
    expo = np.zeros(l + 1, dtype=np.int64)
    v = 1.0
    for k in range(1, l + 1):  # d_{l,0} = prod_k sqrt((2k-1)/(2k))
        v *= np.sqrt((2 * k - 1) / (2.0 * k))
    v *= (-1.0) ** l
    e = 0
    for m in range(l + 1):
        mant[m], expo[m] = v, e
        if m < l:  # d_{l,m+1}/d_{l,m} = -sqrt((l-m)/(l+m+1))
            v *= -np.sqrt((l - m) / (l + m + 1.0))
            if abs(v) < _SCALE_LO:
                v *= _SCALE_HI
                e -= 400
    return mant, expo


def d_halfpi(l):
    """Matrix D[m'+l, m] = d^l_{m'm}(pi/2) for m' = -l..l, m = 0..l."""
    mant, expo = _top_row(l)
    m = np.arange(l + 1, dtype=float)
    out = np.zeros((2 * l + 1, l + 1))
    cur = mant.copy()                  # d_{m',m} scaled by 2^-expo
    nxt = np.zeros(l + 1)              # d_{m'+1,m}
    out[2 * l] = np.ldexp(cur, expo)
    for mp in range(l, 0, -1):         # produce d_{mp-1, m}
        A = np.sqrt((l - mp) * (l + mp + 1.0))
        B = np.sqrt((l + mp) * (l - mp + 1.0))
        new = (2.0 * m * cur - A * nxt) / B
        nxt, cur = cur, new
        big = np.abs(cur) > _SCALE_HI
        if np.any(big):
            cur = np.where(big, cur * _SCALE_LO, cur)
            nxt = np.where(big, nxt * _SCALE_LO, nxt)
            expo = np.where(big, expo + 400, expo)
        out[l + mp - 1] = np.ldexp(cur, expo)
    sign = (-1.0) ** (l + np.arange(l + 1))
    for mp in range(1, l + 1):
        out[l - mp] = sign * out[l + mp]
    return out


def d_beta(l, beta_sign):
    """d^l_{m'm}(beta_sign * pi/2) for m' = -l..l, m = 0..l.

    d_{m'm}(-b) = d_{mm'}(b) = (-1)^{m-m'} d_{m'm}(b).
    """
    D = d_halfpi(l)
    if beta_sign > 0:
        return D
    mp = np.arange(-l, l + 1)[:, None]
    m = np.arange(l + 1)[None, :]
    return D * (-1.0) ** ((m - mp) % 2)


def _indx(ell, m, Nl):
    return (m * (2 * Nl - 1 - m)) // 2 + ell


def rotate_alm_inverse_euler(alm_rot, Nl, alpha, beta_sign, gamma):
    """a from a' = alm of points rotated by R_z(alpha) R_y(+-pi/2) R_z(gamma)."""
    alm_rot = np.asarray(alm_rot, dtype=np.complex128)
    out = np.zeros_like(alm_rot)
    for l in range(Nl):
        D = d_beta(l, beta_sign)
        ms = np.arange(l + 1)
        ap = alm_rot[_indx(l, ms, Nl)]
        full = np.empty(2 * l + 1, dtype=np.complex128)
        full[l:] = ap
        full[:l] = ((-1.0) ** ms[1:] * np.conj(ap[1:]))[::-1]
        mpr = np.arange(-l, l + 1)
        res = (np.exp(1j * mpr * alpha) * full) @ D
        out[_indx(l, ms, Nl)] = res * np.exp(1j * ms * gamma)
    return out


def euler_zyz_matrix(alpha, beta, gamma):
    """Active rotation R_z(alpha) R_y(beta) R_z(gamma) as a 3x3 matrix."""
    def rz(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])

    def ry(b):
        c, s = np.cos(b), np.sin(b)
        return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
    return rz(alpha) @ ry(beta) @ rz(gamma)


def rotate_alm_inverse(alm_rot, Nl, R):
    """Undo the point rotation R for R = R_y(+90 deg) or R_y(-90 deg)."""
    if np.allclose(R, euler_zyz_matrix(0.0, np.pi / 2, 0.0), atol=1e-15):
        return rotate_alm_inverse_euler(alm_rot, Nl, 0.0, +1, 0.0)
    if np.allclose(R, euler_zyz_matrix(0.0, -np.pi / 2, 0.0), atol=1e-15):
        return rotate_alm_inverse_euler(alm_rot, Nl, 0.0, -1, 0.0)
    raise ValueError("rotate_alm_inverse: only R_y(+-90 deg) is supported; "
                     "use rotate_alm_inverse_euler for z-composed rotations")
