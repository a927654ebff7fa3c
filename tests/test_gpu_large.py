"""Large point sets on the GPU.

* A transform of 9e7 points, more than 1280 sort tiles of 65280 points hold:
  the sort takes more tiles rather than larger ones, since its per-tile
  segment counts are 16-bit.  Every point lies in one equatorial segment
  (cos theta in [0.301, 0.305], one bin of the Nx = 64 grid), so each tile
  fills one count to its 65280 points.  FP64 plan, against the sum of three
  transforms of a third of the points each (linearity), within 1e-11 of
  max|a|.
* A batch of point sets (dsht_transform_many, FP32 matrix-core engines)
  whose sets the library groups by their point total: a set of 9e6 points
  goes alone, three sets of 3e6 split into groups of two and one, small
  sets join the last group.  Each set within 1e-6 of max|a| of its single
  FP32 transform.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_one_segment_beyond_sort_tiles(torch_cuda):
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    N = 90_000_000
    g = torch.Generator(device="cuda").manual_seed(7)
    c = 0.301 + 0.004 * torch.rand(N, generator=g, device="cuda", dtype=torch.float64)
    th = torch.acos(c)
    del c
    ph = 2 * math.pi * torch.rand(N, generator=g, device="cuda", dtype=torch.float64)
    w = 0.5 + torch.rand(N, generator=g, device="cuda", dtype=torch.float64)
    sht = DirectSHT(32, 64, dflt_type="float64")
    a, status = sht.transform_device(th, ph, w)
    a = a.clone()
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    parts = torch.zeros_like(a)
    for k in range(3):
        sl = slice(k * N // 3, (k + 1) * N // 3)
        b, _ = sht.transform_device(th[sl], ph[sl], w[sl])
        parts += b
    torch.cuda.synchronize()
    e = rel(a, parts)
    print(f"9e7 points in one segment: {e:.2e}")
    assert e < 1e-11


def test_many_sets_grouped_by_points(torch_cuda):
    torch = torch_cuda
    from directsht_amd.sht import DirectSHT
    sizes = [9_000_000, 3_000_000, 3_000_000, 3_000_000, 1000, 20000, 0, 500]
    g = torch.Generator(device="cuda").manual_seed(8)
    sets = []
    for n in sizes:
        th = torch.acos(2 * torch.rand(n, generator=g, device="cuda", dtype=torch.float64) - 1)
        ph = 2 * math.pi * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
        w = 0.5 + torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
        sets.append((th, ph, w))
    sht = DirectSHT(128, 256, dflt_type="float32")
    alm, status = sht.transform_many_device(sets)
    alm = alm.clone()
    torch.cuda.synchronize()
    assert int(status.max()) == 0
    for i, (th, ph, w) in enumerate(sets):
        if th.numel() == 0:
            assert float(alm[i].abs().max()) == 0.0
            continue
        ref, _ = sht.transform_device(th, ph, w)
        torch.cuda.synchronize()
        e = rel(alm[i], ref)
        print(f"set {i} ({sizes[i]} points): {e:.2e}")
        assert e < 1e-6, i
