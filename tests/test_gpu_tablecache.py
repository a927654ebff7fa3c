"""Plans and 3j tables from the on-disk cache on the GPU: a DirectSHT built
from a cached Y table (dsht_plan_create_from_tables) transforms bit for bit
as a freshly built one (FP64 and FP32 plans, single and weight-set
transforms), its Yv equals the fresh one's, Wigner3j from the cache equals
the built table, and a server started on the cache returns the fresh
plan's alm; one FP64 entry serves both precisions.  Prints the build and
load times of a Nl=512 plan."""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_plan_from_cached_tables_is_bitwise(torch_cuda, tmp_path):
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    g = torch.Generator(device="cuda").manual_seed(5)
    n = 30000
    th = torch.acos(2 * torch.rand(n, generator=g, device="cuda", dtype=torch.float64) - 1)
    ph = 2 * math.pi * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    W = 0.5 + torch.rand((6, n), generator=g, device="cuda", dtype=torch.float64)
    for dt in ("float64", "float32"):
        # the FP64 pass builds and saves the entry; the FP32 plans load it
        # (one FP64 entry serves both precisions)
        fresh = DirectSHT(96, 192, dflt_type=dt)
        first = DirectSHT(96, 192, dflt_type=dt, table_cache=tmp_path)
        cached = DirectSHT(96, 192, dflt_type=dt, table_cache=tmp_path)
        assert first.from_cache == (dt == "float32") and cached.from_cache
        assert np.array_equal(cached.Yv, fresh.Yv)
        a0, _ = fresh.transform_device(th, ph, W[0])
        a0 = a0.clone()
        a1, _ = cached.transform_device(th, ph, W[0])
        torch.cuda.synchronize()
        assert torch.equal(a0, a1), dt
        s0, _ = fresh.transform_sets_device(th, ph, W)
        s0 = s0.clone()
        s1, _ = cached.transform_sets_device(th, ph, W)
        torch.cuda.synchronize()
        assert torch.equal(s0, s1), dt
    # cold start at Nl=512, Nx=1024 (FP32): build against load
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    DirectSHT(512, 1024, dflt_type="float32")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    DirectSHT(512, 1024, dflt_type="float32", table_cache=tmp_path)  # builds and saves
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    DirectSHT(512, 1024, dflt_type="float32", table_cache=tmp_path)  # loads
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    print(f"Nl=512 plan: build {1e3 * (t1 - t0):.1f} ms, build+save {1e3 * (t2 - t1):.1f} ms, "
          f"load {1e3 * (t3 - t2):.1f} ms")


def test_threej_and_server_from_cache(torch_cuda, tmp_path):
    torch = torch_cuda
    from directsht_amd.serve import TransformServer
    from directsht_amd.sht import DirectSHT
    from directsht_amd.threej000 import Wigner3j
    w0 = Wigner3j(40)
    w1 = Wigner3j(40, table_cache=tmp_path)
    w2 = Wigner3j(40, table_cache=tmp_path)
    w3 = Wigner3j.load(40, tmp_path)
    assert not w1.from_cache and w2.from_cache and w3.from_cache
    assert np.array_equal(w0.get_3js(), w2.get_3js())
    assert torch.equal(w0.store_device, w3.store_device)
    rng = np.random.default_rng(9)
    t = np.arccos(rng.uniform(-1, 1, 5000))
    p = rng.uniform(0, 2 * np.pi, 5000)
    w = rng.uniform(0.5, 1.5, 5000)
    DirectSHT(64, 128, dflt_type="float32", table_cache=tmp_path)  # fills the cache
    ref = DirectSHT(64, 128, dflt_type="float32")(t, p, w)
    with TransformServer(max_inflight=2, table_cache=tmp_path) as srv:
        out = srv.submit(t, p, w, Nl=64, Nx=128).result(timeout=120)
        assert srv._tables[(64, 128)].from_cache
    assert np.array_equal(out, ref)
