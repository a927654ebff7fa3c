"""FP32 transforms with weights far from 1 against the FP64 oracle.

The matrix-core scatter splits w h_k(t) e^{-i m phi} into fp16 hi + lo parts;
fp16 covers about 6e-8 .. 6.5e4, so the library scales each transform's
operands by a power of two taken from max |w| and undoes it exactly.  These
tests run weights of 1e-6 (4 pi / N for N = 1e7 points), 1e-3, 1e6 and 1e12
(and one set spanning 1e-6 .. 1e6 with both signs) through the single, weight
set, many-set and served FP32 paths; bar 1e-6 of max|a| against
oracle.refalg.direct_sht (FP64), as for unit weights.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL32 = 1e-6
Nl, Nx = 64, 128
SCALES = (1e-6, 1e-3, 1e6, 1e12)


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def tables():
    from oracle import refalg
    return refalg.make_tables(Nl, Nx, 0.75)


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _points(seed, n):
    rng = np.random.default_rng(seed)
    t = np.arccos(rng.uniform(-1, 1, n))
    p = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    return t, p, w


def _mixed(seed, n):
    """weights of both signs with magnitudes log-uniform over 1e-6 .. 1e6"""
    rng = np.random.default_rng(seed)
    return rng.choice([-1.0, 1.0], n) * np.exp(rng.uniform(np.log(1e-6), np.log(1e6), n))


def test_single_transform_weight_scales(torch_cuda, tables):
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    t, p, w = _points(11, 30000)
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    for scale in SCALES:
        ws = w * scale
        e = rel(sht(t, p, ws), refalg.direct_sht(tables, t, p, ws))
        print(f"single, w x {scale:g}: {e:.2e}")
        assert e < TOL32, scale
    wm = _mixed(12, len(t))
    e = rel(sht(t, p, wm), refalg.direct_sht(tables, t, p, wm))
    print(f"single, mixed 1e-6..1e6: {e:.2e}")
    assert e < TOL32


@pytest.mark.parametrize("engine", ["matrix", "spread"])
def test_weight_set_paths_weight_scales(torch_cuda, tables, engine):
    """one set (the scatter gathers the weights itself) and 5 sets (matrix
    engine: the weight-set scatter; spreading engine: one slot per set, folded
    together), matrix-core contraction"""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    t, p, w = _points(13, 20000)
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    sht.configure(scatter=engine)
    th = torch.from_numpy(t).cuda()
    ph = torch.from_numpy(p).cuda()
    W = np.stack([w * s for s in SCALES] + [_mixed(14, len(t))])
    alm1, _ = sht.transform_sets_device(th, ph, torch.from_numpy(W[:1].copy()).cuda())
    alm, status = sht.transform_sets_device(th, ph, torch.from_numpy(W).cuda())
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    assert rel(alm1[0], refalg.direct_sht(tables, t, p, W[0])) < TOL32
    for s in range(len(W)):
        e = rel(alm[s], refalg.direct_sht(tables, t, p, W[s]))
        print(f"sets, set {s}: {e:.2e}")
        assert e < TOL32, s


def test_many_and_served_weight_scales(torch_cuda, tables):
    torch = torch_cuda
    from directsht_amd.serve import TransformServer
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    data = []
    for k, s in enumerate(SCALES):
        t, p, w = _points(20 + k, 5000 + 3000 * k)
        data.append((t, p, w * s))
    refs = [refalg.direct_sht(tables, *d) for d in data]
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    sets = [tuple(torch.from_numpy(a).cuda() for a in d) for d in data]
    alm, status = sht.transform_many_device(sets)
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    for k in range(len(data)):
        assert rel(alm[k], refs[k]) < TOL32, k
    with TransformServer(max_inflight=2) as srv:
        futs = [srv.submit(*d, Nl=Nl, Nx=Nx) for d in data]
        for k, f in enumerate(futs):
            e = rel(f.result(timeout=120), refs[k])
            print(f"served, w x {SCALES[k]:g}: {e:.2e}")
            assert e < TOL32, k
