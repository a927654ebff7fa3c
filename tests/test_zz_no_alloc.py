"""dsht_transform_noalloc (the allocation-free variant of dsht_transform)
through the C ABI, against dsht_transform on the same inputs.

Two calls in a row on one plan, the second with a smaller input, for FP64
(bar 1e-11 of max|a|) and FP32 (bar 1e-6) plans: the same kernels run in a
different workspace.  The plan's own workspace stays unallocated (its
footprint stays at the tables), and a transform too large for the fixed
buffer is refused with its byte count, alm untouched.
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


def _points(torch, n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    th = torch.acos(2 * torch.rand(n, generator=g, device="cuda", dtype=torch.float64) - 1)
    ph = 2 * math.pi * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    w = 0.5 + torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    return th, ph, w


def _call(fn, plan, th, ph, w, out, status, torch):
    rc = fn(plan, th.numel(), th.data_ptr(), ph.data_ptr(), w.data_ptr(), out.data_ptr(),
            status.data_ptr(), 0, torch.cuda.current_stream().cuda_stream)
    return rc


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("dflt_type,Nl,Nx", [("float64", 64, 128), ("float32", 64, 128),
                                             ("float32", 256, 512)])
def test_noalloc_matches_transform(torch_cuda, dflt_type, Nl, Nx):
    torch = torch_cuda
    from directsht_amd import _lib
    lib = _lib.load()
    prec = _lib.DSHT_FP64 if dflt_type == "float64" else _lib.DSHT_FP32
    plan = ct.c_void_p()
    _lib.check(lib.dsht_plan_create(Nl, Nx, 0.75, prec, ct.byref(plan)), "dsht_plan_create")
    ref_plan = ct.c_void_p()
    _lib.check(lib.dsht_plan_create(Nl, Nx, 0.75, prec, ct.byref(ref_plan)), "dsht_plan_create")
    try:
        fp0 = ct.c_int64()
        _lib.check(lib.dsht_plan_footprint(plan, ct.byref(fp0)), "footprint")
        Nlm = Nl * (Nl + 1) // 2
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        for n, seed in ((200_000, 1), (30_000, 2)):  # second call smaller
            th, ph, w = _points(torch, n, seed)
            a = torch.full((Nlm,), 7.0, dtype=torch.complex128, device="cuda")
            r = torch.empty(Nlm, dtype=torch.complex128, device="cuda")
            rc = _call(lib.dsht_transform_noalloc, plan, th, ph, w, a, status, torch)
            _lib.check(rc, "dsht_transform_noalloc")
            _lib.check(_call(lib.dsht_transform, ref_plan, th, ph, w, r, status, torch),
                       "dsht_transform")
            torch.cuda.synchronize()
            e = _rel(a, r)
            print(f"{dflt_type} Nl={Nl} n={n}: noalloc vs transform {e:.2e}")
            assert e < TOL[dflt_type]
        fp1 = ct.c_int64()
        _lib.check(lib.dsht_plan_footprint(plan, ct.byref(fp1)), "footprint")
        assert fp1.value == fp0.value  # the plan's workspace was never allocated
    finally:
        lib.dsht_plan_destroy(plan)
        lib.dsht_plan_destroy(ref_plan)


def test_noalloc_refuses_oversize(torch_cuda):
    torch = torch_cuda
    from directsht_amd import _lib
    lib = _lib.load()
    plan = ct.c_void_p()
    _lib.check(lib.dsht_plan_create(512, 1024, 0.75, _lib.DSHT_FP32, ct.byref(plan)), "create")
    try:
        th, ph, w = _points(torch, 1000, 3)
        a = torch.full((512 * 513 // 2,), 7.0, dtype=torch.complex128, device="cuda")
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        rc = _call(lib.dsht_transform_noalloc, plan, th, ph, w, a, status, torch)
        torch.cuda.synchronize()
        assert rc == -1  # DSHT_EINVAL
        msg = lib.dsht_last_error().decode()
        assert "needs" in msg and "67108864" in msg, msg
        assert bool((a == 7.0).all())
    finally:
        lib.dsht_plan_destroy(plan)
