"""Batched transforms through the C ABI: many weight sets on shared points
(dsht_transform_sets) and many point sets at one band limit
(dsht_transform_many), against single transforms and the FP64 oracle.

The matrix-core contraction (FP32 plans, 4 or more sets) is an fp16x3 GEMM:
bar 1e-6 of max|a| per set against the single FP32 transform and the FP64
oracle; FP64 plans stay at 1e-11.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"float64": 1e-11, "float32": 1e-6}


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def points(torch, seed, n):
    g = torch.Generator(device="cuda").manual_seed(seed)
    th = torch.acos(2 * torch.rand(n, generator=g, device="cuda", dtype=torch.float64) - 1)
    ph = 2 * math.pi * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    return th, ph


@pytest.mark.parametrize("dflt_type,S,engine", [
    ("float64", 3, None), ("float64", 6, None), ("float32", 3, None), ("float32", 9, None),
    ("float32", 70, None), ("float32", 3, "spread"), ("float32", 21, "spread")])
def test_weight_sets_match_single(torch_cuda, dflt_type, S, engine):
    """Every set of a batch against the single transform of its weights.  On
    the spreading scatter, 21 sets are spread into slots of their own and
    folded in groups of 16 and 5."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    n = 40000
    th, ph = points(torch, 10 + S, n)
    g = torch.Generator(device="cuda").manual_seed(99)
    W = 0.5 + torch.rand((S, n), generator=g, device="cuda", dtype=torch.float64)
    W[0] = 1.0
    W[1] = -W[1] * 1e3  # a set of another sign and scale
    sht = DirectSHT(80, 160, dflt_type=dflt_type)
    if engine:
        sht.configure(scatter=engine)
    alm, status = sht.transform_sets_device(th, ph, W)
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    for s in range(S):
        ref = sht(th, ph, W[s])
        assert rel(alm[s], ref) < TOL[dflt_type], s
    a2 = sht(th.cpu().numpy(), ph.cpu().numpy(), W.cpu().numpy())  # 2-D weights in __call__
    assert a2.shape == (S, sht.Nlm) and rel(a2, alm) < TOL[dflt_type]


@pytest.mark.parametrize("dflt_type,S", [("float64", 5), ("float32", 2), ("float32", 7)])
def test_many_point_sets_match_single(torch_cuda, dflt_type, S):
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    rng = np.random.default_rng(S)
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(500), np.log(1e5), S))]
    sizes[-1] = 0  # an empty set
    sets = []
    for i, n in enumerate(sizes):
        th, ph = points(torch, 100 + i, n)
        w = 0.5 + torch.rand(n, device="cuda", dtype=torch.float64)
        sets.append((th, ph, w))
    sht = DirectSHT(64, 128, dflt_type=dflt_type)
    alm, status = sht.transform_many_device(sets)
    torch.cuda.synchronize()
    for i, (th, ph, w) in enumerate(sets):
        if th.numel() == 0:
            assert float(alm[i].abs().max()) == 0.0
            continue
        ref = sht(th, ph, w)
        assert rel(alm[i], ref) < TOL[dflt_type], i


def test_weight_sets_vs_oracle(torch_cuda):
    torch = torch_cuda
    from oracle import refalg
    from directsht_amd.sht import DirectSHT
    n = 6000
    th, ph = points(torch, 3, n)
    W = 0.5 + torch.rand((5, n), device="cuda", dtype=torch.float64)
    tables = refalg.make_tables(40, 80, 0.75)
    for dflt_type in ("float64", "float32"):
        sht = DirectSHT(40, 80, dflt_type=dflt_type)
        alm, _ = sht.transform_sets_device(th, ph, W)
        torch.cuda.synchronize()
        for s in range(5):
            ref = refalg.direct_sht(tables, th.cpu().numpy(), ph.cpu().numpy(), W[s].cpu().numpy())
            assert rel(alm[s], ref) < TOL[dflt_type], (dflt_type, s)


def test_weight_sets_lognormal_c3_like(torch_cuda):
    """Shape of the many-weight-set workload at reduced size: constant plus
    lognormal-modulated weight sets on uniform points; every set within 1e-6
    of the FP64 plan."""
    torch = torch_cuda
    from directsht_amd import mocks
    from directsht_amd.sht import DirectSHT
    n, S = 200000, 16
    th, ph = points(torch, 4, n)
    W = mocks.lognormal_weight_sets(th, ph, S, lmax=64, ntheta=128, nphi=256)
    s32 = DirectSHT(192, 384, dflt_type="float32")
    s64 = DirectSHT(192, 384, dflt_type="float64")
    a32, _ = s32.transform_sets_device(th, ph, W)
    a64, _ = s64.transform_sets_device(th, ph, W)
    torch.cuda.synchronize()
    for s in range(S):
        assert rel(a32[s], a64[s]) < 1e-6, s


def test_many_point_sets_in_groups(torch_cuda):
    """40 point sets at one band limit: three groups of up to 16 sets sorted,
    scattered and folded together (FP32, matrix-core engines), with empty sets
    between others, a set with NaN points (dropped), a polar-only and an
    equatorial-only set; each set within 1e-6 of its single FP32 transform
    and of the FP64 oracle."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    rng = np.random.default_rng(40)
    S, Nl = 40, 96
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(50), np.log(2e5), S))]
    sizes[5] = sizes[17] = sizes[39] = 0
    sets = []
    for i, n in enumerate(sizes):
        if i == 8:  # polar only
            c = torch.empty(n, device="cuda", dtype=torch.float64).uniform_(0.8, 1.0)
            th = torch.acos(c * torch.where(torch.rand(n, device="cuda") < 0.5, -1.0, 1.0))
            ph = 2 * math.pi * torch.rand(n, device="cuda", dtype=torch.float64)
        elif i == 9:  # equatorial only
            th = torch.acos(torch.empty(n, device="cuda", dtype=torch.float64).uniform_(-0.7, 0.7))
            ph = 2 * math.pi * torch.rand(n, device="cuda", dtype=torch.float64)
        else:
            th, ph = points(torch, 500 + i, n)
        if i == 20 and n > 3:
            th = th.clone()
            th[::3] = float("nan")
        w = 0.5 + torch.rand(n, device="cuda", dtype=torch.float64)
        sets.append((th, ph, w))
    sht = DirectSHT(Nl, 2 * Nl, dflt_type="float32")
    alm, status = sht.transform_many_device(sets)
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    tables = refalg.make_tables(Nl, 2 * Nl, 0.75)
    for i, (th, ph, w) in enumerate(sets):
        if th.numel() == 0:
            assert float(alm[i].abs().max()) == 0.0, i
            continue
        assert rel(alm[i], sht(th, ph, w)) < TOL["float32"], i
        if i % 7 == 0 or i in (8, 9, 20):
            t, p, ww = (a.cpu().numpy() for a in (th, ph, w))
            keep = ~np.isnan(t)
            ref = refalg.direct_sht(tables, t[keep], p[keep], ww[keep])
            assert rel(alm[i], ref) < TOL["float32"], i


def test_many_point_sets_dense_and_sparse(torch_cuda):
    """A point-set batch on the automatic engine: the sets with at least 256
    points per segment go one by one through the spreading scatter, the
    others in groups of consecutive sparse sets through the matrix engine,
    all into one matrix-core contraction.  Dense and sparse sets interleaved
    (and an empty one), within a reservation: each set within 1e-6 of its
    single transform and of the FP64 oracle, and nothing grows."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx = 48, 96
    dense_n = 256 * 4 * (Nx - 1)
    sizes = [dense_n + 5000, 700, 20000, 0, 3 * dense_n, 150, dense_n, 45000, 9000, 2 * dense_n - 1]
    sets = []
    for i, n in enumerate(sizes):
        th, ph = points(torch, 900 + i, n)
        sets.append((th, ph, 0.5 + torch.rand(n, device="cuda", dtype=torch.float64)))
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    assert [sht.scatter_engine(n) for n in sizes].count("spread") == 4
    sht.reserve(max(sizes), len(sizes), weight_sets=False)
    f0 = sht.footprint()
    alm, status = sht.transform_many_device(sets)
    again, _ = sht.transform_many_device(sets)
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    assert sht.footprint() == f0
    assert torch.equal(alm, again)
    tables = refalg.make_tables(Nl, Nx, 0.75)
    for i, (th, ph, w) in enumerate(sets):
        if th.numel() == 0:
            assert float(alm[i].abs().max()) == 0.0, i
            continue
        assert rel(alm[i], sht(th, ph, w)) < TOL["float32"], i
        if i in (0, 1, 4, 7):
            ref = refalg.direct_sht(tables, *(a.cpu().numpy() for a in (th, ph, w)))
            assert rel(alm[i], ref) < TOL["float32"], i


@pytest.mark.parametrize("S", [3, 5, 21])
def test_weight_sets_node_spread(torch_cuda, S):
    """Weight sets on the spreading scatter with 512 node pairs (its node
    variant): from 4 sets each set is spread and folded in one launch into
    its slot of the matrix-core contraction (3 sets: into the plain records,
    contracted set by set).  Sets of other signs and scales (x -1e3, x 1e-6,
    all zero), and (21 sets) points with NaN theta, which the binning drops:
    each set within 1e-6 of its single transform, of the FP64 plan, and two
    of them of the FP64 oracle; a second call bit for bit equal."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx, n = 96, 1024, 30000
    th, ph = points(torch, 500 + S, n)
    if S == 21:
        th[::97] = float("nan")
    g = torch.Generator(device="cuda").manual_seed(7 + S)
    W = 0.5 + torch.rand((S, n), generator=g, device="cuda", dtype=torch.float64)
    W[1] = -W[1] * 1e3
    W[2] = W[2] * 1e-6
    if S > 3:
        W[3] = 0.0
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    sht.configure(scatter="spread")
    alm, status = sht.transform_sets_device(th, ph, W)
    again, _ = sht.transform_sets_device(th, ph, W)
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    assert torch.equal(alm, again)
    f64 = DirectSHT(Nl, Nx, dflt_type="float64")
    for s in range(S):
        if s == 3:
            assert float(alm[s].abs().max()) == 0.0
            continue
        assert rel(alm[s], sht(th, ph, W[s])) < TOL["float32"], s
        assert rel(alm[s], f64(th, ph, W[s])) < TOL["float32"], s
    tables = refalg.make_tables(Nl, Nx, 0.75)
    t, p = th.cpu().numpy(), ph.cpu().numpy()
    keep = ~np.isnan(t)
    for s in (1, 2):
        ref = refalg.direct_sht(tables, t[keep], p[keep], W[s].cpu().numpy()[keep])
        assert rel(alm[s], ref) < TOL["float32"], s


def test_many_point_sets_dense_node_spread(torch_cuda):
    """A point-set batch at 512 node pairs on the automatic engine: the dense
    sets (256 or more points per segment) are spread and folded in one launch
    each into their slots (the node variant), the sparse ones go in groups
    through the matrix engine.  Each set within 1e-6 of the FP64 plan's
    transform of it, one dense and two sparse ones also of the FP64 oracle."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx = 40, 1024
    dense_n = 256 * 4 * (Nx - 1)
    sizes = [dense_n + 777, 900, 0, 2 * dense_n - 3, 12000, dense_n]
    sets = []
    for i, n in enumerate(sizes):
        th, ph = points(torch, 950 + i, n)
        w = 0.5 + torch.rand(n, device="cuda", dtype=torch.float64)
        if i == 3:
            w = -w * 1e4
        sets.append((th, ph, w))
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    assert [sht.scatter_engine(n) for n in sizes].count("spread") == 3
    alm, status = sht.transform_many_device(sets)
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    f64 = DirectSHT(Nl, Nx, dflt_type="float64")
    tables = refalg.make_tables(Nl, Nx, 0.75)
    for i, (th, ph, w) in enumerate(sets):
        if th.numel() == 0:
            assert float(alm[i].abs().max()) == 0.0, i
            continue
        assert rel(alm[i], f64(th, ph, w)) < TOL["float32"], i
        if i in (0, 1, 4):
            ref = refalg.direct_sht(tables, *(a.cpu().numpy() for a in (th, ph, w)))
            assert rel(alm[i], ref) < TOL["float32"], i


@pytest.mark.parametrize("Nx,engine", [(192, None), (1024, "spread")])
def test_reserve_kinds_cover_their_transforms(torch_cuda, Nx, engine):
    """A point-set reservation holds every dsht_transform_many group of its
    sizes, a weight-set reservation every dsht_transform_sets: the plan's
    footprint does not change over transforms within them (also on the node
    spreading variant, whose weight sets need sorted weight rows)."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    n, S = 50000, 20
    rng = np.random.default_rng(21)
    sets = []
    for i, m in enumerate(int(x) for x in rng.integers(1, n + 1, S)):
        th, ph = points(torch, 700 + i, m)
        sets.append((th, ph, 0.5 + torch.rand(m, device="cuda", dtype=torch.float64)))
    base = DirectSHT(96, Nx, dflt_type="float32")
    if engine:
        base.configure(scatter=engine)
    sp = base.clone()
    sp.reserve(n, S, weight_sets=False)
    f0 = sp.footprint()
    sp.transform_many_device(sets)
    sp.transform_many_device(sets[:5])
    torch.cuda.synchronize()
    assert sp.footprint() == f0
    sw = base.clone()
    sw.reserve(n, S, point_sets=False)
    f1 = sw.footprint()
    th, ph = points(torch, 800, n)
    W = 0.5 + torch.rand((S, n), device="cuda", dtype=torch.float64)
    sw.transform_sets_device(th, ph, W)
    sw.transform_sets_device(th[:1000], ph[:1000], W[:7, :1000].contiguous())
    torch.cuda.synchronize()
    assert sw.footprint() == f1
