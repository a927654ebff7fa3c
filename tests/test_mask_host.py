"""Host-side steps of MaskDeconvolution (sht/mask_deconvolution.py): binning,
bin-bin inversion, decoupling and normalization, with the mode-coupling matrix
supplied by the oracle (no GPU).  The decoupling of a pseudo-spectrum made
from a C_l that is constant within each bin returns those constants exactly,
which pins the inversion and the order of the products."""
import numpy as np
import pytest

from oracle import coupling


@pytest.fixture(scope="module")
def md_cls():
    from directsht_amd.mask_deconvolution import MaskDeconvolution
    return MaskDeconvolution


def make(md_cls, Nl, W):
    md = md_cls(Nl, W, verbose=False, precompute_3j_and_Mll=False)
    md.Mll = coupling.mode_coupling(Nl, W)
    return md


def test_binning_matrix(md_cls):
    md = md_cls(100, np.ones(5), verbose=False, precompute_3j_and_Mll=False)
    b = md.binning_matrix("linear", 0, 16)
    assert b.shape == (6, 100)
    np.testing.assert_allclose(b.sum(axis=1), 1.0)
    for i in range(6):
        np.testing.assert_array_equal(np.nonzero(b[i])[0], np.arange(16 * i, 16 * i + 16))
    b = md.binning_matrix("linear", 4, 32)
    assert b.shape == (3, 100) and b[0, 4] == 1 / 32 and b[0, 3] == 0 and b[2, 99] == 1 / 32
    s = md.binning_matrix("sqrt", 2, 3)
    lo = 2
    for row in s:
        w = int(np.ceil(np.sqrt(4.0 * lo) + 3))
        np.testing.assert_array_equal(np.nonzero(row)[0], np.arange(lo, lo + w))
        lo += w
    assert lo + int(np.ceil(np.sqrt(4.0 * lo) + 3)) > 100
    assert md.binning_matrix("linear", 0, 200).shape == (0, 100)
    with pytest.raises(RuntimeError):
        md.binning_matrix("log")


def test_padding_and_window(md_cls):
    md = md_cls(10, np.arange(4.0), verbose=False, precompute_3j_and_Mll=False)
    assert md.W_l.size == 19 and (md.W_l[4:] == 0).all()
    assert md.W(3) == 3.0 and md.W(0, debug=True) == 4 * np.pi and md.W(5, debug=True) == 0
    md = md_cls(10, np.arange(30.0), verbose=False, precompute_3j_and_Mll=False)
    assert md.W_l.size == 30


def test_decoupling_recovers_bin_constants(md_cls):
    Nl = 40
    rng = np.random.default_rng(2)
    W = np.exp(-np.arange(2 * Nl - 1) / 6.0) * 3.0
    md = make(md_cls, Nl, W)
    bins = md.binning_matrix("linear", 0, 8)
    c = rng.uniform(1, 2, bins.shape[0])
    C = ((bins > 0).astype(float) * c[:, None]).sum(axis=0)   # constant within each bin
    pseudo = md.Mll @ C
    ells, dec = md(pseudo, bins)
    np.testing.assert_allclose(dec, c, rtol=1e-12)
    np.testing.assert_allclose(ells, [3.5 + 8 * i for i in range(5)])
    # window matrix maps the same theory to the same bandpowers
    _, conv = md.convolve_theory_Cls(C, bins)
    np.testing.assert_allclose(conv, c, rtol=1e-12)
    # Cl shorter than Nl is zero-padded, longer is cut
    _, d1 = md(pseudo[:30], bins)
    _, d2 = md(np.concatenate([pseudo[:30], np.zeros(10)]), bins)
    np.testing.assert_array_equal(d1, d2)
    _, d3 = md(np.concatenate([pseudo, np.ones(7)]), bins)
    np.testing.assert_array_equal(d3, dec)
    with pytest.raises(AssertionError):
        md.convolve_theory_Cls(C[:20], bins)
    with pytest.raises(RuntimeError):
        md(pseudo, bins, mode="other")


def test_normalization(md_cls):
    Nl = 16
    md = make(md_cls, Nl, np.array([4 * np.pi]))        # full sky: factor 1, M = identity
    assert md.compute_normalization_factor() == pytest.approx(1.0, rel=1e-15)
    np.testing.assert_allclose(md.Mll, np.eye(Nl), atol=1e-14)
    W = np.array([2.0, 0.5, 0.25])
    md = make(md_cls, Nl, W)
    want = 4 * np.pi / (2.0 + 3 * 0.5 + 5 * 0.25)
    assert md.compute_normalization_factor() == pytest.approx(want, rel=1e-15)
    bins = md.binning_matrix("linear", 0, 4)
    Cl = np.linspace(1, 2, Nl)
    _, nb = md(Cl, bins, mode="normalization")
    np.testing.assert_allclose(nb, want * (bins @ Cl), rtol=1e-15)
    np.testing.assert_allclose(md.window_matrix(bins, "normalization"), want * (bins @ md.Mll),
                               rtol=1e-15)
