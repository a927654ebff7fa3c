"""FP64 restatement of the reference's Wigner-3j (m1 = m2 = m3 = 0) table and
mode-coupling matrix.

TEST INFRASTRUCTURE ONLY -- never imported by the product path.

  * threej000: the recursion of sht/threej000.py:9-42 (and
    csht/threej000_helper.c:28-56) for one symbol;
  * threej000_table: the packed table of sht/threej000.py:107-116
    (Wigner3j.get_3js, with the closed-form entries of fill_simple3j,
    :46-55) or, with prefill=False, csht/threej000_helper.c:62-77
    (make_table, pure recursion);
  * tj_index: sht/threej000.py:61-75;
  * mode_coupling: MaskDeconvolution.get_M, sht/mask_deconvolution.py:171-192,
    with the window padding of :33-34 and the full-sky debug window of
    :155-169.

Parity pins (tests/test_oracle_coupling.py):
  * threej000_table(prefill=False) equals the reference's compiled C
    make_table (oracle/_ref/threej000_helper.so) bit for bit, and the
    prefilled table equals it to an ulp;
  * the symbols the reference's own test compares with sympy
    (sht/test_3j.py:20-40) against sympy's exact wigner_3j, and every symbol
    with j1 < 12 against it;
  * the Racah closed form (l1 l2 l3; 0 0 0)^2 = (g!)^2/(2g+1)! *
    prod_i (2g - 2l_i)!/((g - l_i)!)^2 (exact rationals), and the orthogonality
    sum_l3 (2 l3 + 1) (l1 l2 l3; 0 0 0)^2 = 1;
  * mode_coupling: the identity for the full-sky debug window (the
    reference's own check, sht/mask_deconvolution.py:162-167), the closed form
    M[l1, l2] = (2 l2 + 1) c / 4pi for a constant window, linearity in the
    window, and the reciprocity (2 l1 + 1) M[l1, l2] = (2 l2 + 1) M[l2, l1].
"""
import numpy as np

__all__ = ["threej000", "tj_index", "threej000_table", "dense_threej000", "mode_coupling"]


def tj_index(l1, l2, l3):
    """Packed index of (l1 l2 l3; 0 0 0) -- sht/threej000.py:61-75: with the
    l's sorted j1 >= j2 >= j3, j1(j1+1)(j1+2)/6 + j2(j2+1)/2 + j3."""
    j1 = max(l1, l2, l3)
    j3 = min(l1, l2, l3)
    j2 = l1 + l2 + l3 - j1 - j3
    return (j1 * (j1 + 1) * (j1 + 2)) // 6 + (j2 * (j2 + 1)) // 2 + j3


def _prefill_value(j2):
    """(j2+1 j2 1; 0 0 0) as fill_simple3j writes it (sht/threej000.py:51-55)."""
    return -((-1.0) ** j2) * np.sqrt((j2 + 1.0) / (2 * j2 + 1.0) / (2 * j2 + 3.0))


def threej000(l1, l2, l3, prefill=False):
    """One symbol by the reference's recursion (sht/threej000.py:9-42,
    csht/threej000_helper.c:28-56): sort j1 >= j2 >= j3; zero for odd J;
    at j3 = 0 the value is (-1)^j1/sqrt(2 j1 + 1) if j1 = j2, else zero;
    otherwise fac * (j1, j2+1, j3-1) with
    fac = sqrt((J-2j2-1)(J-2j3+2) / ((J-2j2)(J-2j3+1))).  The product is taken
    from the terminal value outwards, as the recursive return does.  With
    prefill, the chain also stops at (j+1, j, 1), whose value the Python
    class pre-fills in closed form."""
    J = l1 + l2 + l3
    if J % 2:
        return 0.0
    facs = []
    a, b, c = l1, l2, l3
    while True:
        j1 = max(a, b, c)
        j3 = min(a, b, c)
        j2 = J - j1 - j3
        if j3 == 0:
            val = ((-1.0) ** j1) / np.sqrt(2 * j1 + 1.0) if j1 == j2 else 0.0
            break
        if prefill and j3 == 1 and j1 == j2 + 1:
            val = _prefill_value(j2)
            break
        num = (J - 2 * j2 - 1) * (J - 2 * j3 + 2)
        den = (J - 2 * j2) * (J - 2 * j3 + 1)
        facs.append(np.sqrt(float(num) / float(den)))
        a, b, c = j1, j2 + 1, j3 - 1
    for f in reversed(facs):
        val = f * val
    return val


def _table_indices(N):
    j1 = np.concatenate([np.full((j + 1) * (j + 2) // 2, j) for j in range(N)])
    j2 = np.concatenate([np.repeat(np.arange(j + 1), np.arange(1, j + 2)) for j in range(N)])
    j3 = np.concatenate([np.concatenate([np.arange(k + 1) for k in range(j + 1)]) for j in range(N)])
    return j1, j2, j3


def threej000_table(N, prefill=True):
    """Packed table of (j1 j2 j3; 0 0 0) for N > j1 >= j2 >= j3 >= 0,
    N(N+1)(N+2)/6 values at tj_index -- Wigner3j(N).store
    (sht/threej000.py:107-116; prefill=True, the fill_simple3j entries in
    closed form) or make_table (csht/threej000_helper.c:62-77;
    prefill=False).  Vectorised: every entry walks its recursion chain in
    lock step, the factors are kept per step and multiplied into the terminal
    value from the last step back, as threej000 does."""
    j1, j2, j3 = _table_indices(N)
    n = j1.size
    J = j1 + j2 + j3
    out = np.zeros(n)
    live = (J % 2 == 0)
    a, b, c = j1.copy(), j2.copy(), j3.copy()
    term = np.zeros(n)
    steps = []            # (indices still walking, their factor) per step
    idx = np.nonzero(live)[0]
    while idx.size:
        s = np.sort(np.stack([a[idx], b[idx], c[idx]]), axis=0)
        k1, k2, k3 = s[2], s[1], s[0]
        Ji = J[idx]
        end0 = k3 == 0
        term[idx[end0]] = np.where(k1[end0] == k2[end0],
                                   ((-1.0) ** k1[end0]) / np.sqrt(2 * k1[end0] + 1.0), 0.0)
        done = end0
        if prefill:
            endp = (k3 == 1) & (k1 == k2 + 1)
            term[idx[endp]] = -((-1.0) ** k2[endp]) * np.sqrt(
                (k2[endp] + 1.0) / (2 * k2[endp] + 1.0) / (2 * k2[endp] + 3.0))
            done = done | endp
        go = ~done
        gi = idx[go]
        if gi.size:
            num = (Ji[go] - 2 * k2[go] - 1) * (Ji[go] - 2 * k3[go] + 2)
            den = (Ji[go] - 2 * k2[go]) * (Ji[go] - 2 * k3[go] + 1)
            steps.append((gi, np.sqrt(num.astype(np.float64) / den.astype(np.float64))))
            a[gi], b[gi], c[gi] = k1[go], k2[go] + 1, k3[go] - 1
        idx = gi
    val = term
    for gi, f in reversed(steps):
        val[gi] = f * val[gi]
    out[live] = val[live]
    return out


def dense_threej000(N, table=None):
    """(l1 l2 l3; 0 0 0) for 0 <= l1, l2, l3 < N as a dense N^3 array, from
    the packed table (the symbol is symmetric in its arguments for m = 0)."""
    if table is None:
        table = threej000_table(N)
    l = np.arange(N)
    L1, L2, L3 = np.meshgrid(l, l, l, indexing="ij")
    S = np.sort(np.stack([L1, L2, L3]), axis=0)
    k3, k2, k1 = S[0], S[1], S[2]
    return table[(k1 * (k1 + 1) * (k1 + 2)) // 6 + (k2 * (k2 + 1)) // 2 + k3]


def mode_coupling(Nl, W_l, debug=False):
    """MaskDeconvolution.get_M (sht/mask_deconvolution.py:171-192):
        M[l1, l2] = (2 l2 + 1)/(4 pi) * sum_{l3 = |l1-l2|}^{l1+l2}, l1+l2+l3 even
                    (2 l3 + 1) W(l3) (l1 l2 l3; 0 0 0)^2,
    summed in ascending l3, then times (2 l2 + 1), then divided by 4 pi.
    W_l is zero-padded to 2 Nl - 1 values (:33-34); debug replaces it with
    the full-sky window W(0) = 4 pi, W(l > 0) = 0 (:162-167).  The 3j's are
    the Wigner3j(2 Nl - 1) table (:42)."""
    W = np.asarray(W_l, dtype=np.float64)
    W = np.pad(W, (0, max(0, 2 * Nl - 1 - W.size)))
    if debug:
        W = np.zeros(2 * Nl - 1)
        W[0] = 4 * np.pi
    N3 = 2 * Nl - 1
    T = dense_threej000(N3)[:Nl, :Nl, :]
    M = np.zeros((Nl, Nl))
    for l3 in range(N3):                    # ascending l3; zero symbols add nothing
        M += (2 * l3 + 1) * W[l3] * T[:, :, l3] ** 2
    M *= (2 * np.arange(Nl) + 1)[None, :]
    M /= 4 * np.pi
    return M
