"""On-disk table cache (directsht_amd/tablecache.py), on the CPU: entries
round-trip exactly, a missing key reads as None, a damaged or mismatched entry
is refused, and Wigner3j.load serves the packed 3j table from the cache
(values checked against oracle.coupling.threej000)."""
import json
import os

import numpy as np
import pytest

from directsht_amd.tablecache import TableCache


def test_ylm_round_trip_and_keys(tmp_path):
    tc = TableCache(tmp_path)
    Nl, Nx, xmax = 9, 7, 0.75
    Yv = np.random.default_rng(0).standard_normal((Nl * (Nl + 1) // 2, Nx))
    assert tc.load_ylm(Nl, Nx, xmax) is None
    tc.save_ylm(Nl, Nx, xmax, Yv)
    back = tc.load_ylm(Nl, Nx, xmax)
    assert back.dtype == np.float64 and np.array_equal(back, Yv)
    # other keys are other entries
    assert tc.load_ylm(Nl, Nx, 0.7) is None
    assert tc.load_ylm(Nl, Nx + 1, xmax) is None
    assert tc.load_ylm(Nl + 1, Nx, xmax) is None
    with pytest.raises(ValueError):
        tc.save_ylm(Nl, Nx, xmax, Yv[:-1])  # wrong shape


def test_damaged_entries_are_refused(tmp_path):
    tc = TableCache(tmp_path)
    Nl, Nx, xmax = 6, 5, 0.75
    Yv = np.arange(Nl * (Nl + 1) // 2 * Nx, dtype=np.float64).reshape(-1, Nx)
    stem = tc.save_ylm(Nl, Nx, xmax, Yv)
    # one flipped byte in the data
    with open(stem + ".npy", "r+b") as f:
        f.seek(-3, os.SEEK_END)
        b = f.read(1)
        f.seek(-3, os.SEEK_END)
        f.write(bytes([b[0] ^ 0x10]))
    with pytest.raises(ValueError, match="checksum"):
        tc.load_ylm(Nl, Nx, xmax)
    # a header naming another key
    tc.save_ylm(Nl, Nx, xmax, Yv)
    with open(stem + ".json") as f:
        head = json.load(f)
    head["key"]["Nx"] = Nx + 1
    with open(stem + ".json", "w") as f:
        json.dump(head, f)
    with pytest.raises(ValueError, match="header"):
        tc.load_ylm(Nl, Nx, xmax)
    # no temporary files left behind
    assert not [n for n in os.listdir(tmp_path) if n.endswith(".tmp")]


def test_wigner3j_load_from_cache(tmp_path):
    from directsht_amd.threej000 import Wigner3j
    from oracle import coupling
    N = 12
    store = coupling.threej000_table(N)
    tc = TableCache(tmp_path)
    with pytest.raises(KeyError):
        Wigner3j.load(N, tc, device="cpu")
    tc.save_threej(N, store)
    w = Wigner3j.load(N, tc, device="cpu")
    assert w.from_cache and np.array_equal(w.get_3js(), store)
    for l1, l2, l3 in ((2, 3, 5), (4, 4, 4), (11, 6, 5), (7, 2, 5)):
        assert w(l1, l2, l3) == pytest.approx(coupling.threej000(l1, l2, l3), abs=1e-15)
