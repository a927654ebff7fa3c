"""Angular power spectra on the device (dsht_alm2cl through DirectSHT.alm2cl
and DirectSHT.cross_spectra) against the FP64 oracle (refalg.alm2cl,
refalg.alm2cl_cross): FP64 sums in the same order of m, so 1e-12."""
import numpy as np
import pytest

from oracle import refalg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rand_alms(rng, S, Nl):
    Nlm = Nl * (Nl + 1) // 2
    a = rng.normal(size=(S, Nlm)) + 1j * rng.normal(size=(S, Nlm))
    ell, m = refalg.getlm(Nl - 1)
    a[:, m == 0] = a[:, m == 0].real
    return a


@pytest.mark.parametrize("Nl,S", [(2, 1), (7, 5), (97, 1), (97, 64), (300, 3)])
def test_alm2cl_vs_oracle(torch_cuda, Nl, S):
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    sht = DirectSHT(Nl, 8)
    rng = np.random.default_rng(Nl + S)
    a = rand_alms(rng, S, Nl)
    cl = sht.alm2cl(a if S > 1 else a[0])                   # numpy in -> numpy out
    assert isinstance(cl, np.ndarray) and cl.shape == ((S, Nl) if S > 1 else (Nl,))
    cl = cl.reshape(S, Nl)
    for s in range(S):
        np.testing.assert_allclose(cl[s], refalg.alm2cl(a[s], Nl), rtol=1e-12, atol=1e-300)
    cd = sht.alm2cl(torch.from_numpy(a).cuda())             # device in -> device out
    assert cd.is_cuda and cd.shape == (S, Nl)
    np.testing.assert_allclose(cd.cpu().numpy(), cl, rtol=1e-12)


def test_alm2cl_cabi_edges(torch_cuda):
    """Nl = 1 (monopole only), a leading dimension above Nlm, and the argument checks."""
    torch = torch_cuda
    import ctypes as ct
    from directsht_amd import _lib
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    a = torch.tensor([[3.0 + 7.0j, 0], [-2.0 + 1.0j, 0]], dtype=torch.complex128, device="cuda")
    cl = torch.full((3, 1), -1.0, dtype=torch.float64, device="cuda")
    assert lib.dsht_alm2cl(1, 2, a.data_ptr(), 2, cl.data_ptr(), 1, stream) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(cl.cpu().numpy()[:, 0], [9.0, -6.0, 4.0])  # m = 0: real parts
    Nl, ld, S = 9, 50, 3
    rng = np.random.default_rng(1)
    ah = np.zeros((S, ld), complex)
    ah[:, :45] = rand_alms(rng, S, Nl)
    ah[:, 45:] = np.nan                                     # padding must not be read
    ad = torch.from_numpy(ah).cuda()
    cl = torch.empty((S, Nl), dtype=torch.float64, device="cuda")
    assert lib.dsht_alm2cl(Nl, S, ad.data_ptr(), ld, cl.data_ptr(), 0, stream) == 0
    torch.cuda.synchronize()
    for s in range(S):
        np.testing.assert_allclose(cl[s].cpu().numpy(), refalg.alm2cl(ah[s, :45], Nl), rtol=1e-12)
    for args in [(0, 1, ad.data_ptr(), ld), (Nl, 0, ad.data_ptr(), ld),
                 (Nl, 65, ad.data_ptr(), ld), (Nl, 1, ad.data_ptr(), 44), (Nl, 1, None, ld)]:
        assert lib.dsht_alm2cl(*args, cl.data_ptr(), 0, stream) != 0
        assert b"dsht_alm2cl" in lib.dsht_last_error()


@pytest.mark.parametrize("Nl,S", [(2, 1), (33, 4), (97, 64)])
def test_cross_spectra_vs_oracle(torch_cuda, Nl, S):
    from directsht_amd.sht import DirectSHT
    sht = DirectSHT(Nl, 8)
    rng = np.random.default_rng(3 * Nl + S)
    a = rand_alms(rng, S, Nl)
    cx = sht.cross_spectra(a)
    assert cx.shape == (S * (S + 1) // 2, Nl)
    p = 0
    for s in range(S):
        for t in range(s, S):
            np.testing.assert_allclose(cx[p], refalg.alm2cl_cross(a[s], a[t], Nl), rtol=1e-12,
                                       atol=1e-12 * np.abs(cx[p]).max())
            p += 1
    auto = sht.alm2cl(a)
    diag = [s * S - s * (s - 1) // 2 for s in range(S)]  # (s, s) in row order
    np.testing.assert_allclose(cx[diag], auto.reshape(S, Nl), rtol=1e-12)


def test_spectrum_of_a_transform(torch_cuda):
    """alm2cl of a transform's output, as the reference's callers use it."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    rng = np.random.default_rng(9)
    n = 20000
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    sht = DirectSHT(64, 128)
    alm = sht(th, ph, w)
    np.testing.assert_allclose(sht.alm2cl(alm), refalg.alm2cl(alm, 64), rtol=1e-12)
