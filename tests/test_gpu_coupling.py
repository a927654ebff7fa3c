"""Wigner-3j table and mode-coupling matrix on the device (dsht_threej000_table,
dsht_mode_coupling through Wigner3j and MaskDeconvolution) against the oracle
(oracle/coupling.py, itself pinned to the reference's C table and sympy).
FP64 throughout: 1e-12 relative to the largest entry; the reference's
recursion itself is accurate to ~1e-14 at these sizes."""
import ctypes as ct

import numpy as np
import pytest

from oracle import coupling, cref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("N", [1, 2, 7, 64, 150])
def test_threej_table_vs_oracle(torch_cuda, N):
    from directsht_amd.threej000 import Wigner3j
    w = Wigner3j(N)
    got = w.store
    assert got.shape == (N * (N + 1) * (N + 2) // 6,)
    ref = coupling.threej000_table(N)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-300)
    assert (got[ref == 0] == 0).all() and (np.sign(got) == np.sign(ref)).all()
    if cref.threej_available() and N <= 64:
        np.testing.assert_allclose(got, cref.threej000_table(N), rtol=1e-13, atol=1e-300)
    if N >= 3:
        assert w(2, 1, 1) == got[coupling.tj_index(2, 1, 1)] == w(1, 2, 1)


def test_threej_table_large(torch_cuda):
    """Wigner3j(1127), the table MaskDeconvolution(564) builds
    (sht/mask_deconvolution.py:42): sampled entries vs the recursion,
    the reference test's points (sht/test_3j.py:20-40) and orthogonality."""
    torch = torch_cuda
    from directsht_amd.threej000 import threej000_table_device
    N = 1127
    tab = threej000_table_device(N)
    rng = np.random.default_rng(11)
    S = np.sort(rng.integers(0, N, size=(400, 3)), axis=1)[:, ::-1]
    idx = [coupling.tj_index(*s) for s in S]
    got = tab[torch.tensor(idx, device=tab.device)].cpu().numpy()
    want = np.array([coupling.threej000(*map(int, s), prefill=True) for s in S])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-300)
    for l1, l2, l3 in [(999, 998, 1), (998, 997, 5), (2, 6, 10), (8, 12, 16), (20, 24, 28)]:
        np.testing.assert_allclose(float(tab[coupling.tj_index(l1, l2, l3)]),
                                   coupling.threej000(l1, l2, l3, prefill=True), rtol=1e-12)
    # sum_l3 (2 l3 + 1) (l1 l2 l3; 0 0 0)^2 = 1 (l1 + l2 < N)
    for l1, l2 in [(0, 0), (300, 7), (560, 560), (1000, 120)]:
        l3 = np.arange(abs(l1 - l2), l1 + l2 + 1)
        ii = [coupling.tj_index(l1, l2, int(k)) for k in l3]
        v = tab[torch.tensor(ii, device=tab.device)].cpu().numpy()
        np.testing.assert_allclose(np.sum((2 * l3 + 1) * v * v), 1.0, rtol=1e-12)


@pytest.mark.parametrize("Nl,nW,signed", [(1, 1, False), (2, 3, False), (17, 10, True),
                                          (64, 127, False), (100, 400, True)])
def test_mode_coupling_vs_oracle(torch_cuda, Nl, nW, signed):
    from directsht_amd.mask_deconvolution import MaskDeconvolution
    rng = np.random.default_rng(Nl)
    W = rng.uniform(-1 if signed else 0, 1, nW)
    md = MaskDeconvolution(Nl, W, verbose=False)
    ref = coupling.mode_coupling(Nl, W)
    np.testing.assert_allclose(md.Mll, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    if not signed:
        np.testing.assert_allclose(md.Mll, ref, rtol=1e-12)
    np.testing.assert_allclose(md.get_M(debug=True), np.eye(Nl), atol=1e-13)


@pytest.mark.parametrize("Nl", [564, 2048])
def test_mode_coupling_identities_large(torch_cuda, Nl):
    """Full-sky debug window -> identity (sht/mask_deconvolution.py:162-167);
    constant window -> (2 l2 + 1) c / 4pi; linearity; reciprocity; a few
    entries against the oracle's recursion."""
    from directsht_amd.mask_deconvolution import mode_coupling_device
    I = mode_coupling_device(Nl, np.array([4 * np.pi])).cpu().numpy()
    np.testing.assert_allclose(I, np.eye(Nl), atol=1e-12)
    c = 0.75
    M = mode_coupling_device(Nl, np.full(2 * Nl - 1, c)).cpu().numpy()
    want = (2 * np.arange(Nl) + 1) * c / (4 * np.pi)
    np.testing.assert_allclose(M, np.broadcast_to(want, M.shape), rtol=1e-12)
    rng = np.random.default_rng(Nl)
    W1 = np.exp(-np.arange(2 * Nl - 1) / 40.0) * rng.uniform(0.5, 1.5, 2 * Nl - 1)
    W2 = rng.uniform(0, 1, 300)
    M1 = mode_coupling_device(Nl, W1).cpu().numpy()
    M2 = mode_coupling_device(Nl, W2).cpu().numpy()
    W12 = 2 * W1
    W12[:300] -= 3 * W2
    M12 = mode_coupling_device(Nl, W12).cpu().numpy()
    np.testing.assert_allclose(M12, 2 * M1 - 3 * M2, atol=1e-12 * np.abs(M12).max())
    ell = (2 * np.arange(Nl) + 1)[:, None]
    np.testing.assert_allclose(ell * M1, (ell * M1).T, rtol=1e-13)
    for l1, l2 in [(0, 5), (Nl - 1, Nl - 1), (Nl // 2, 40)]:
        s = 0.0
        for l3 in range(abs(l1 - l2), l1 + l2 + 1, 2):
            s += (2 * l3 + 1) * W1[l3] * coupling.threej000(l1, l2, l3, prefill=True) ** 2
        np.testing.assert_allclose(M1[l1, l2], s * (2 * l2 + 1) / (4 * np.pi), rtol=1e-12)


def test_mask_deconvolution_end_to_end(torch_cuda):
    """Bandpowers of a pseudo-spectrum made from a C_l constant within bins
    come back as those constants; the window matrix gives the same."""
    from directsht_amd.mask_deconvolution import MaskDeconvolution
    Nl = 96
    W = 3.0 * np.exp(-np.arange(2 * Nl - 1) / 10.0)
    md = MaskDeconvolution(Nl, W, verbose=False)
    np.testing.assert_allclose(md.Mll, coupling.mode_coupling(Nl, W), rtol=1e-12)
    bins = md.binning_matrix("linear", 0, 16)
    c = np.linspace(1.0, 2.0, bins.shape[0])
    C = ((bins > 0).astype(float) * c[:, None]).sum(axis=0)
    _, dec = md(md.Mll @ C, bins)
    np.testing.assert_allclose(dec, c, rtol=1e-11)
    _, conv = md.convolve_theory_Cls(C, bins)
    np.testing.assert_allclose(conv, c, rtol=1e-11)
    # lazily built table of the reference's attribute
    assert md.w3j000.Nl == 2 * Nl - 1
    np.testing.assert_allclose(md.w3j000(40, 30, 20), coupling.threej000(40, 30, 20, prefill=True),
                               rtol=1e-13)


def test_cabi_argument_errors(torch_cuda):
    torch = torch_cuda
    from directsht_amd import _lib
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    buf = torch.zeros(64, dtype=torch.float64, device="cuda")
    for n in (0, 4097):
        assert lib.dsht_threej000_table(n, buf.data_ptr(), st) != 0
        assert b"dsht_threej000_table" in lib.dsht_last_error()
    assert lib.dsht_threej000_table(3, None, st) != 0
    assert lib.dsht_mode_coupling(0, buf.data_ptr(), 1, buf.data_ptr(), st) != 0
    assert lib.dsht_mode_coupling(4097, buf.data_ptr(), 1, buf.data_ptr(), st) != 0
    assert lib.dsht_mode_coupling(4, buf.data_ptr(), -1, buf.data_ptr(), st) != 0
    assert lib.dsht_mode_coupling(4, None, 3, buf.data_ptr(), st) != 0
    assert lib.dsht_mode_coupling(4, buf.data_ptr(), 3, None, st) != 0
    assert b"dsht_mode_coupling" in lib.dsht_last_error()
    # nW = 0: zero window, zero matrix
    M = torch.full((4, 4), 7.0, dtype=torch.float64, device="cuda")
    assert lib.dsht_mode_coupling(4, None, 0, M.data_ptr(), st) == 0
    torch.cuda.synchronize()
    assert (M == 0).all()


def test_mode_coupling_beyond_lds_range(torch_cuda):
    """Nl above 2048 takes the kernel variant that gathers from global memory
    instead of LDS: the same identities, and entries against the recursion."""
    from directsht_amd.mask_deconvolution import mode_coupling_device
    Nl = 2500
    I = mode_coupling_device(Nl, np.array([4 * np.pi])).cpu().numpy()
    np.testing.assert_allclose(I, np.eye(Nl), atol=1e-12)
    c = 0.3
    M = mode_coupling_device(Nl, np.full(2 * Nl - 1, c)).cpu().numpy()
    np.testing.assert_allclose(M, np.broadcast_to((2 * np.arange(Nl) + 1) * c / (4 * np.pi), M.shape),
                               rtol=1e-12)
    rng = np.random.default_rng(3)
    W = rng.uniform(0, 1, 700)
    M = mode_coupling_device(Nl, W).cpu().numpy()
    for l1, l2 in [(2499, 2), (300, 250), (1800, 1600)]:
        s = 0.0
        for l3 in range(abs(l1 - l2), min(l1 + l2, 699) + 1, 2):
            s += (2 * l3 + 1) * W[l3] * coupling.threej000(l1, l2, l3, prefill=True) ** 2
        np.testing.assert_allclose(M[l1, l2], s * (2 * l2 + 1) / (4 * np.pi), rtol=1e-11,
                                   atol=1e-14 * np.abs(M).max())
