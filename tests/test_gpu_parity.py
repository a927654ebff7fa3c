"""GPU parity of the HIP path (libdsht.so via DirectSHT) against the FP64
oracle (oracle.refalg), the exact direct sum and the csht golden fixtures.

Tolerances: FP64 mode within 1e-11 and FP32 mode within 1e-6 of
the FP64 restatement of the reference algorithm, measured as
max|a - a_ref| / max|a_ref| (and the rms analogue).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"float64": 1e-11, "float32": 1e-6}


def rel_max(a, r):
    return float(np.max(np.abs(a - r)) / np.max(np.abs(r)))


def rel_rms(a, r):
    return float(np.linalg.norm(a - r) / np.linalg.norm(r))


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _uniform(rng, n, lo=-1.0, hi=1.0):
    return (np.arccos(rng.uniform(lo, hi, n)), rng.uniform(0, 2 * np.pi, n),
            rng.uniform(0.5, 1.5, n))


_TABLES = {}


def oracle_tables(Nl, Nx, xmax=0.75):
    from oracle import refalg
    key = (Nl, Nx, xmax)
    if key not in _TABLES:
        _TABLES[key] = refalg.make_tables(Nl, Nx, xmax)
    return _TABLES[key]


def test_tables_match_oracle(torch_cuda):
    from directsht_amd.sht import DirectSHT
    for Nl, Nx in ((64, 128), (200, 301)):
        s = DirectSHT(Nl, Nx)
        _, Yv, Yd = oracle_tables(Nl, Nx)
        assert np.array_equal(s.x, oracle_tables(Nl, Nx)[0])
        for mine, ref in ((s.Yv, Yv), (s.Yd, Yd)):
            scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
            assert np.max(np.abs(mine - ref) / scale) < 1e-14
        exact = np.mean(s.Yv == Yv)
        print(f"Nl={Nl} Nx={Nx}: bit-identical Yv entries {exact:.4f}")


@pytest.mark.parametrize("dflt_type", ["float64", "float32"])
@pytest.mark.parametrize("Nl,Nx,n", [(64, 128, 5000), (100, 201, 20000), (256, 512, 100000)])
def test_uniform_sphere_vs_oracle(torch_cuda, dflt_type, Nl, Nx, n):
    """Config-1-like: uniform sphere (25% polar), weights U[0.5,1.5]."""
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    rng = np.random.default_rng(1000 + Nl)
    t, p, w = _uniform(rng, n)
    ref = refalg.direct_sht(oracle_tables(Nl, Nx), t, p, w)
    a = DirectSHT(Nl, Nx, dflt_type=dflt_type)(t, p, w)
    assert a.dtype == np.complex128 and a.shape == ref.shape
    e, r = rel_max(a, ref), rel_rms(a, ref)
    print(f"{dflt_type} Nl={Nl} n={n}: max {e:.2e} rms {r:.2e}")
    assert e < TOL[dflt_type] and r < TOL[dflt_type]


def test_csht_golden(torch_cuda, golden_dir):
    from directsht_amd.sht import DirectSHT
    d = np.load(f"{golden_dir}/csht_alm_equ.npz")
    a = DirectSHT(int(d["Nl_rand"]), int(d["Nx_rand"]))(d["theta"], d["phi"], d["wt"])
    assert rel_max(a, d["alm_rand"]) < 1e-12
    a = DirectSHT(int(d["Nl_det"]), int(d["Nx_det"]))(d["theta_det"], d["phi_det"], d["wt_det"])
    assert rel_max(a, d["alm_det"]) < 1e-12


def test_nodes_exact_vs_direct_sum(torch_cuda):
    from directsht_amd.sht import DirectSHT
    from oracle import exact
    Nl, Nx = 96, 192
    xx = oracle_tables(Nl, Nx)[0]
    rng = np.random.default_rng(7)
    n = 4000
    t = np.arccos(rng.choice([-1.0, 1.0], n) * xx[rng.integers(0, Nx - 1, n)])
    p = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    a = DirectSHT(Nl, Nx)(t, p, w)
    assert rel_max(a, exact.exact_alm(Nl, t, p, w)) < 1e-12


def test_polar_only_and_equatorial_only_and_one_hemisphere(torch_cuda):
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx = 80, 160
    s = DirectSHT(Nl, Nx)
    T = oracle_tables(Nl, Nx)
    rng = np.random.default_rng(8)
    for lo, hi in ((0.8, 1.0), (-1.0, -0.76), (-0.7, 0.7), (0.0, 0.74), (-0.74, -0.01)):
        t, p, w = _uniform(rng, 3000, lo, hi)
        assert rel_max(s(t, p, w), refalg.direct_sht(T, t, p, w)) < 1e-11, (lo, hi)


def test_edge_cases(torch_cuda):
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx = 32, 65  # odd Nx: a padding node in the last pair
    s = DirectSHT(Nl, Nx)
    T = oracle_tables(Nl, Nx)
    z = s(np.array([]), np.array([]), np.array([]))
    assert z.shape == (Nl * (Nl + 1) // 2,) and np.all(z == 0)
    # single points: equator, x on a node, polar, exactly at the poles
    for th in (np.pi / 2, np.arccos(T[0][7]), 0.2, 0.0, np.pi):
        a = s(np.array([th]), np.array([1.3]), np.array([2.0]))
        r = refalg.direct_sht(T, np.array([th]), np.array([1.3]), np.array([2.0]))
        assert np.max(np.abs(a - r)) <= 1e-12 * np.max(np.abs(r)), th
    # NaN theta is dropped, as the reference's masks drop it
    rng = np.random.default_rng(9)
    t, p, w = _uniform(rng, 500)
    tn = t.copy()
    tn[::7] = np.nan
    keep = ~np.isnan(tn)
    assert rel_max(s(tn, p, w), refalg.direct_sht(T, t[keep], p[keep], w[keep])) < 1e-12
    # length mismatch
    with pytest.raises(AssertionError):
        s(t, p[:-1], w)
    # equatorial theta outside [0, pi] fails the basic_sht range check
    with pytest.raises(AssertionError):
        s(np.array([-1.2, 1.0]), np.array([0.0, 0.0]), np.ones(2))


def test_small_xmax_polar_failure_returns_zeros(torch_cuda, capsys):
    from directsht_amd.sht import DirectSHT
    s = DirectSHT(16, 32, xmax=0.6)
    a = s(np.array([0.9, 1.5, 2.0]), np.array([0.0, 1.0, 2.0]), np.ones(3))
    assert np.all(a == 0)
    assert "sin(pi/4)" in capsys.readouterr().out


def test_basic_sht_semantics(torch_cuda):
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx = 48, 96
    s = DirectSHT(Nl, Nx)
    T = oracle_tables(Nl, Nx)
    rng = np.random.default_rng(10)
    t, p, w = _uniform(rng, 2000, -0.74, 0.74)
    assert rel_max(s.basic_sht(t, p, w, reg_factor=4.0), refalg.basic_sht(T, t, p, w)) < 1e-12
    with pytest.raises(ValueError):
        s.basic_sht(np.array([]), np.array([]), np.array([]))
    with pytest.raises(AssertionError):
        s.basic_sht(np.array([0.1]), np.array([0.0]), np.array([1.0]))


def test_device_tensor_inputs_and_repeat(torch_cuda):
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx = 128, 256
    s = DirectSHT(Nl, Nx, dflt_type="float32")
    rng = np.random.default_rng(11)
    t, p, w = _uniform(rng, 50000)
    ref = refalg.direct_sht(oracle_tables(Nl, Nx), t, p, w)
    td, pd, wd = (torch.from_numpy(a).cuda() for a in (t, p, w))
    a1 = s(td, pd, wd)
    assert isinstance(a1, torch.Tensor) and a1.is_cuda and a1.dtype == torch.complex128
    a1 = a1.cpu().numpy()
    a2 = s(t, p, w)  # same plan, second call, numpy path
    assert rel_max(a1, ref) < 1e-6 and rel_max(a2, ref) < 1e-6
    # the order inside a segment is the input order and every reduction runs
    # in a fixed order: two calls give the same bits
    assert np.array_equal(a1, a2)


def test_csht_c_abi(torch_cuda, golden_dir):
    """make_table / do_transform through the C ABI, as csht/sht.py binds them."""
    import ctypes as ct
    from directsht_amd import _lib
    lib = _lib.load()
    dp = ct.POINTER(ct.c_double)
    d = np.load(f"{golden_dir}/csht_tables_small.npz")
    Nl, Nx = int(d["Nl"]), int(d["Nx"])
    Yv = np.zeros(Nl * (Nl + 1) // 2 * Nx)
    Yd = np.zeros_like(Yv)
    assert lib.make_table(Nl, Nx, 0.75, Yv.ctypes.data_as(dp), Yd.ctypes.data_as(dp)) == 0
    np.testing.assert_allclose(Yv.reshape(-1, Nx), d["Yv"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(Yd.reshape(-1, Nx), d["Yd"], rtol=1e-12, atol=1e-13)
    g = np.load(f"{golden_dir}/csht_alm_equ.npz")
    Nl, Nx = int(g["Nl_rand"]), int(g["Nx_rand"])
    Nlm = Nl * (Nl + 1) // 2
    Yv = np.zeros(Nlm * Nx)
    Yd = np.zeros_like(Yv)
    assert lib.make_table(Nl, Nx, 0.75, Yv.ctypes.data_as(dp), Yd.ctypes.data_as(dp)) == 0
    cost = np.cos(g["theta"])
    alm = np.zeros(Nlm, dtype=np.complex128)
    for sel, sign in ((cost < 0, -1), (cost >= 0, 1)):
        c = np.ascontiguousarray(cost[sel])
        ph = np.ascontiguousarray(g["phi"][sel])
        wt = np.ascontiguousarray(g["wt"][sel])
        carr, sarr = np.zeros(Nlm), np.zeros(Nlm)
        rc = lib.do_transform(Nl, Nx, 0.75, Yv.ctypes.data_as(dp), Yd.ctypes.data_as(dp), len(c),
                              c.ctypes.data_as(dp), ph.ctypes.data_as(dp), wt.ctypes.data_as(dp),
                              carr.ctypes.data_as(dp), sarr.ctypes.data_as(dp))
        assert rc == 0
        a = carr - 1j * sarr
        if sign < 0:
            ell = np.concatenate([np.arange(m, Nl) for m in range(Nl)])
            m = np.concatenate([np.full(Nl - m, m) for m in range(Nl)])
            a = a * (-1.0) ** (ell - m)
        alm += a
    assert rel_max(alm, g["alm_rand"]) < 1e-12


def test_scatter_engines_agree(torch_cuda):
    """FP32 plans: matrix-core (fp16x3 MFMA) and vector (FP32 VALU) scatters
    against the FP64 GPU path, which is pinned to the oracle above."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    Nl, Nx = 300, 600
    rng = np.random.default_rng(12)
    t, p, w = (torch.from_numpy(a).cuda() for a in _uniform(rng, 60000))
    ref = DirectSHT(Nl, Nx, dflt_type="float64")(t, p, w).cpu().numpy()
    for eng in ("matrix", "vector"):
        s = DirectSHT(Nl, Nx, dflt_type="float32")
        s.configure(scatter=eng)
        a = s(t, p, w).cpu().numpy()
        e, r = rel_max(a, ref), rel_rms(a, ref)
        print(f"{eng}: max {e:.2e} rms {r:.2e}")
        assert e < 1e-6 and r < 1e-6, eng


@pytest.mark.parametrize("Nl", [40, 256, 257, 512, 513])
def test_matrix_scatter_tile_widths(torch_cuda, Nl):
    """The matrix-core scatter's m tile (256, 512 or 1024 orders, chosen by Nl)
    and the spreading scatter's grid (n = 4 Nl rounded up to a power of two;
    odd and even log2 n) at and beside each width: FP32 matrix, vector and
    spread engines against the FP64 path."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    Nx = 2 * Nl
    rng = np.random.default_rng(Nl)
    t, p, w = (torch.from_numpy(a).cuda() for a in _uniform(rng, 30000))
    ref = DirectSHT(Nl, Nx, dflt_type="float64")(t, p, w).cpu().numpy()
    for eng in ("matrix", "vector", "spread"):
        s = DirectSHT(Nl, Nx, dflt_type="float32")
        s.configure(scatter=eng)
        a = s(t, p, w).cpu().numpy()
        print(f"{eng}: max {rel_max(a, ref):.2e} rms {rel_rms(a, ref):.2e}")
        assert rel_max(a, ref) < 1e-6 and rel_rms(a, ref) < 1e-6, eng


@pytest.mark.parametrize("Nl", [1100, 1536])
def test_high_l_second_m_tile(torch_cuda, Nl):
    """Nl > 1024 uses a second m tile (matrix scatter) / more lanes (vector)."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    rng = np.random.default_rng(Nl)
    t, p, w = (torch.from_numpy(a).cuda() for a in _uniform(rng, 20000))
    ref = DirectSHT(Nl, 2 * Nl, dflt_type="float64")(t, p, w).cpu().numpy()
    a = DirectSHT(Nl, 2 * Nl, dflt_type="float32")(t, p, w).cpu().numpy()
    assert rel_max(a, ref) < 1e-6 and rel_rms(a, ref) < 1e-6
    s = DirectSHT(Nl, 2 * Nl, dflt_type="float32")
    s.configure(scatter="spread")  # grid n = 8192 (128 KB of LDS)
    a = s(t, p, w).cpu().numpy()
    assert rel_max(a, ref) < 1e-6 and rel_rms(a, ref) < 1e-6


def test_auto_engine_and_repeat(torch_cuda):
    """The default ('auto') engine takes the spreading scatter once a transform
    averages 256 points per segment (4 (Nx-1) segments) on grids too small for
    its node variant, and the matrix engine below; both against the oracle at
    <= 1e-6, and the spreading path (64-bit fixed-point grids) gives the same
    bits on every call."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    from oracle import refalg
    Nl, Nx = 64, 128
    nseg = 4 * (Nx - 1)
    s = DirectSHT(Nl, Nx, dflt_type="float32")
    assert s.scatter_engine(256 * nseg - 1) == "matrix"
    assert s.scatter_engine(256 * nseg) == "spread"
    tabs = oracle_tables(Nl, Nx)
    for n in (50000, 256 * nseg + 1000):
        rng = np.random.default_rng(n)
        t, p, w = _uniform(rng, n)
        ref = refalg.direct_sht(tabs, t, p, w)
        td, pd, wd = (torch.from_numpy(a).cuda() for a in (t, p, w))
        a1 = s(td, pd, wd).cpu().numpy()
        a2 = s(td, pd, wd).cpu().numpy()
        e = rel_max(a1, ref)
        print(f"n={n} ({s.scatter_engine(n)}): max {e:.2e}")
        assert e < 1e-6 and rel_rms(a1, ref) < 1e-6
        assert np.array_equal(a1, a2)


@pytest.mark.parametrize("Nx", [1024, 1023])
def test_node_spread_variant(torch_cuda, Nx):
    """The node variant of the spreading scatter (grids of 512 or more node
    pairs, Nl <= 1024: scatter and fold in one kernel) is taken from 64 points
    per segment; against the FP64 path at <= 1e-6 and the same bits on a
    repeat, with an odd Nx (the last node of the last pair is padding), points
    on both branches and both hemispheres, and points exactly at the grid's
    first and last nodes."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    Nl = 512
    nseg = 4 * (Nx - 1)
    s = DirectSHT(Nl, Nx, dflt_type="float32")
    assert s.scatter_engine(64 * nseg - 1) == "matrix"
    assert s.scatter_engine(64 * nseg) == "spread"
    rng = np.random.default_rng(Nx)
    t, p, w = _uniform(rng, 64 * nseg + 5000)
    xg = np.arange(Nx) * 0.75 / (Nx - 1)
    t[:4] = np.arccos([xg[0], -xg[0], xg[-1], -xg[-1]])  # at the first and last nodes
    td, pd, wd = (torch.from_numpy(a).cuda() for a in (t, p, w))
    ref = DirectSHT(Nl, Nx, dflt_type="float64")(td, pd, wd).cpu().numpy()
    a1 = s(td, pd, wd).cpu().numpy()
    a2 = s(td, pd, wd).cpu().numpy()
    e, r = rel_max(a1, ref), rel_rms(a1, ref)
    print(f"node variant Nx={Nx}: max {e:.2e} rms {r:.2e}")
    assert e < 1e-6 and r < 1e-6
    assert np.array_equal(a1, a2)


def test_node_spread_variant_inputs(torch_cuda):
    """The node variant on the inputs the small-grid tests give the other
    engines: polar-only, equatorial-only and one-hemisphere point sets (units
    with no points skip their transform), NaN points (dropped), weights far
    from 1 and of both signs (the fixed-point scale follows max |w|), each
    against the FP64 path at <= 1e-6."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    Nl, Nx = 512, 1024
    n = 64 * 4 * (Nx - 1) + 2000
    s32 = DirectSHT(Nl, Nx, dflt_type="float32")
    s64 = DirectSHT(Nl, Nx, dflt_type="float64")
    assert s32.scatter_engine(n) == "spread"
    rng = np.random.default_rng(77)
    cases = []
    for lo, hi in ((0.8, 1.0), (-1.0, -0.76), (-0.7, 0.7), (0.0, 0.74), (-0.74, -0.01)):
        cases.append(((lo, hi), _uniform(rng, n, lo, hi)))
    t, p, w = _uniform(rng, n)
    t[::97] = np.nan
    cases.append(("nan", (t, p, w)))
    t, p, w = _uniform(rng, n)
    for scale in (1e-6, 1e12):
        cases.append((f"w x {scale:g}", (t, p, w * scale)))
    ws = w * np.exp(rng.uniform(np.log(1e-6), np.log(1e6), n)) * rng.choice([-1.0, 1.0], n)
    cases.append(("mixed weights", (t, p, ws)))
    # 250000 points at one position: every point of a node on the same 8 grid
    # cells (the per-node fixed-point scale keeps the cell sums in range)
    t, p, w = _uniform(rng, n)
    t[:250000], p[:250000] = 1.1, 2.0
    cases.append(("coincident points", (t, p, w)))
    for name, (t, p, w) in cases:
        td, pd, wd = (torch.from_numpy(a).cuda() for a in (t, p, w))
        ref = s64(td, pd, wd).cpu().numpy()
        a = s32(td, pd, wd).cpu().numpy()
        e = rel_max(a, ref)
        print(f"node variant, {name}: max {e:.2e}")
        assert e < 1e-6, name


@pytest.mark.parametrize("n", [1000, 100000])
def test_single_ring_accumulation(torch_cuda, n):
    """All points in one spline bin (a ring of constant theta): the longest
    possible runs, the hardest case for FP32 accumulation."""
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    Nl, Nx = 128, 256
    rng = np.random.default_rng(n)
    t = np.full(n, np.arccos(0.4123))
    p = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    ref = DirectSHT(Nl, Nx, dflt_type="float64")(t, p, w)
    for eng in ("matrix", "vector", "spread"):
        s = DirectSHT(Nl, Nx, dflt_type="float32")
        s.configure(scatter=eng)
        a = s(t, p, w)
        assert rel_max(a, ref) < 1e-6 and rel_rms(a, ref) < 1e-6, eng
