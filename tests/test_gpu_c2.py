"""The headline configuration (Nl=1000, Nx=2048, lognormal cap points) and a
band limit above 2048, on the GPU.

* csht golden vectors (the reference's own C, tests/golden/csht_c2_equ.npz):
  the equatorial part of points drawn from the cap generator, FP64 plan within
  1e-11 and FP32 plan within 1e-6 of max|a| (alm at every 37th index).
* Full size (1e7 cap points, about 36 % polar): properties that fix the
  result independently of its size -- FP32 against FP64 within 1e-6 (max and
  rms), the reflection theta -> pi - theta giving (-1)^(l-m) a_lm (exact for
  the spline approximant in both branches: the equatorial one folds |cos|,
  and in the rotated frame the reflection is a rotation by pi about x' --
  parity and conjugation), within 1e-11 (FP64) and 1e-6 (FP32), and
  linearity in the weights within 1e-6 (FP32).
* Nl=2100 (tests/golden/csht_nl2100.npz): make_table and do_transform
  through the C ABI within 1e-12, DirectSHT plans within 1e-11 / 1e-6, and
  the polar branch (rotation with 67 KB of LDS per degree) by the reflection
  property and FP32 against FP64.
"""
import ctypes as ct
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


def rel_max(a, r):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    r = r.cpu().numpy() if hasattr(r, "cpu") else np.asarray(r)
    return float(np.max(np.abs(a - r)) / np.max(np.abs(r)))


def rel_rms(a, r):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    r = r.cpu().numpy() if hasattr(r, "cpu") else np.asarray(r)
    return float(np.linalg.norm(a - r) / np.linalg.norm(r))


def parity(Nl):
    ell = np.concatenate([np.arange(m, Nl) for m in range(Nl)])
    m = np.concatenate([np.full(Nl - m, m) for m in range(Nl)])
    return (-1.0) ** (ell - m)


@pytest.mark.parametrize("dflt_type", ["float64", "float32"])
def test_c2_csht_golden(torch_cuda, golden_dir, dflt_type):
    from directsht_amd.sht import DirectSHT
    d = np.load(f"{golden_dir}/csht_c2_equ.npz")
    s = DirectSHT(int(d["Nl"]), int(d["Nx"]), dflt_type=dflt_type)
    a = s(d["theta"], d["phi"], d["wt"])[d["aidx"]]
    e, r = rel_max(a, d["alm"]), rel_rms(a, d["alm"])
    print(f"C2 csht golden, {dflt_type}: max {e:.2e} rms {r:.2e} ({len(d['theta'])} points)")
    assert e < TOL[dflt_type] and r < TOL[dflt_type]


def test_c2_full_size_properties(torch_cuda):
    torch = torch_cuda
    from directsht_amd import mocks
    from directsht_amd.sht import DirectSHT
    Nl, Nx, N = 1000, 2048, 10_000_000
    th, ph, w = mocks.lognormal_cap_points(N, seed=2)
    npol = int((torch.cos(th).abs() >= 0.75).sum())
    print(f"C2: {N} points, {npol / N:.3f} polar")
    sgn = torch.from_numpy(parity(Nl)).cuda()
    s64 = DirectSHT(Nl, Nx, dflt_type="float64")
    a64, _ = s64.transform_device(th, ph, w)
    r64, _ = s64.transform_device(math.pi - th, ph, w)
    torch.cuda.synchronize()
    e = rel_max(r64, sgn * a64)
    print(f"C2 FP64 reflection: {e:.2e}")
    assert e < TOL["float64"]
    a64 = a64.clone()
    del s64, r64
    torch.cuda.empty_cache()
    s32 = DirectSHT(Nl, Nx, dflt_type="float32")
    # the default engine at this size: the spreading scatter (fixed-point
    # grids, so the same bits on every call), then the recurrence contraction
    assert s32.scatter_engine(N) == "spread"
    a32, _ = s32.transform_device(th, ph, w)
    a32 = a32.clone()
    again, _ = s32.transform_device(th, ph, w)
    r32, _ = s32.transform_device(math.pi - th, ph, w)
    r32 = r32.clone()
    torch.cuda.synchronize()
    assert torch.equal(again, a32), "two calls differ"
    e, r = rel_max(a32, a64), rel_rms(a32, a64)
    print(f"C2 FP32 vs FP64: max {e:.2e} rms {r:.2e}")
    assert e < TOL["float32"] and r < TOL["float32"]
    sm = DirectSHT(Nl, Nx, dflt_type="float32")
    sm.configure(scatter="matrix")
    am, _ = sm.transform_device(th, ph, w)
    torch.cuda.synchronize()
    e = rel_max(am, a64)
    print(f"C2 FP32 matrix engine vs FP64: max {e:.2e}")
    assert e < TOL["float32"]
    del sm, am, again
    torch.cuda.empty_cache()
    e = rel_max(r32, sgn * a64)
    print(f"C2 FP32 reflection: {e:.2e}")
    assert e < TOL["float32"]
    w2 = mocks.lognormal_weights(th, ph, seed=302)
    b1 = a32
    b2, _ = s32.transform_device(th, ph, w2)
    b2 = b2.clone()
    b3, _ = s32.transform_device(th, ph, w + 2.0 * w2)
    torch.cuda.synchronize()
    e = rel_max(b3, b1 + 2.0 * b2)
    print(f"C2 FP32 linearity: {e:.2e}")
    assert e < TOL["float32"]


def test_nl2100_c_abi_golden(torch_cuda, golden_dir):
    from directsht_amd import _lib
    lib = _lib.load()
    dp = ct.POINTER(ct.c_double)
    d = np.load(f"{golden_dir}/csht_nl2100.npz")
    Nl, Nx = int(d["Nl"]), int(d["Nx"])
    Nlm = Nl * (Nl + 1) // 2
    Yv = np.zeros(Nlm * Nx)
    Yd = np.zeros_like(Yv)
    assert lib.make_table(Nl, Nx, 0.75, Yv.ctypes.data_as(dp), Yd.ctypes.data_as(dp)) == 0
    Yv2, Yd2 = Yv.reshape(Nlm, Nx), Yd.reshape(Nlm, Nx)
    for mine, ref in ((Yv2[d["ridx"]], d["Yv"]), (Yd2[d["ridx"]], d["Yd"])):
        scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
        assert np.max(np.abs(mine - ref) / scale) < 1e-12
    cost = np.cos(d["theta"])
    alm = np.zeros(Nlm, dtype=np.complex128)
    sg = parity(Nl)
    for sel, neg in ((cost < 0, True), (cost >= 0, False)):
        c = np.ascontiguousarray(cost[sel])
        ph = np.ascontiguousarray(d["phi"][sel])
        wt = np.ascontiguousarray(d["wt"][sel])
        carr, sarr = np.zeros(Nlm), np.zeros(Nlm)
        rc = lib.do_transform(Nl, Nx, 0.75, Yv.ctypes.data_as(dp), Yd.ctypes.data_as(dp), len(c),
                              c.ctypes.data_as(dp), ph.ctypes.data_as(dp), wt.ctypes.data_as(dp),
                              carr.ctypes.data_as(dp), sarr.ctypes.data_as(dp))
        assert rc == 0
        alm += (carr - 1j * sarr) * (sg if neg else 1.0)
    e = rel_max(alm[d["aidx"]], d["alm"])
    print(f"Nl=2100 do_transform vs csht: {e:.2e}")
    assert e < 1e-12
    # beyond the transform band limit: a clear error, arrays untouched
    carr = np.full(10, 7.0)
    Nbig = 4097
    rc = lib.do_transform(Nbig, 8, 0.75, Yv.ctypes.data_as(dp), Yd.ctypes.data_as(dp), 1,
                          cost[:1].ctypes.data_as(dp), cost[:1].ctypes.data_as(dp),
                          cost[:1].ctypes.data_as(dp), carr.ctypes.data_as(dp),
                          carr.ctypes.data_as(dp))
    assert rc == 1 and np.all(carr == 7.0)
    assert b"4096" in lib.dsht_last_error()


def test_nl2100_plans(torch_cuda, golden_dir):
    torch = torch_cuda
    from directsht_amd import mocks
    from directsht_amd.sht import DirectSHT
    d = np.load(f"{golden_dir}/csht_nl2100.npz")
    Nl, Nx = int(d["Nl"]), int(d["Nx"])
    for dt in ("float64", "float32"):
        a = DirectSHT(Nl, Nx, dflt_type=dt)(d["theta"], d["phi"], d["wt"])[d["aidx"]]
        e = rel_max(a, d["alm"])
        print(f"Nl=2100 {dt} vs csht: {e:.2e}")
        assert e < TOL[dt]
    # polar branch: uniform sphere, 25 % polar
    th, ph, w = mocks.uniform_sphere(200000, seed=21)
    sgn = torch.from_numpy(parity(Nl)).cuda()
    s64 = DirectSHT(Nl, 2 * Nl, dflt_type="float64")
    a64, _ = s64.transform_device(th, ph, w)
    a64 = a64.clone()
    r64, _ = s64.transform_device(math.pi - th, ph, w)
    torch.cuda.synchronize()
    e = rel_max(r64, sgn * a64)
    print(f"Nl=2100 FP64 reflection: {e:.2e}")
    assert e < TOL["float64"]
    s32 = DirectSHT(Nl, 2 * Nl, dflt_type="float32")
    a32, _ = s32.transform_device(th, ph, w)
    torch.cuda.synchronize()
    e = rel_max(a32, a64)
    print(f"Nl=2100 FP32 vs FP64: {e:.2e}")
    assert e < TOL["float32"]
