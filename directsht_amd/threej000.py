"""Wigner 3j symbols (l1 l2 l3; 0 0 0) -- drop-in for the reference's
``sht.threej000.Wigner3j`` (sht/threej000.py:77-116).

The packed table is built on the device by ``dsht_threej000_table`` (closed
form of the m = 0 symbol, HIP kernel in libdsht.so) and kept there;
``store`` copies it to the host on first use.
"""
import numpy as np
import torch

from . import _lib


def get_index(l1, l2, l3):
    """Packed index of (l1 l2 l3; 0 0 0) (sht/threej000.py:61-75): with the
    l's sorted j1 >= j2 >= j3, j1(j1+1)(j1+2)/6 + j2(j2+1)/2 + j3."""
    j1 = max(l1, l2, l3)
    j3 = min(l1, l2, l3)
    j2 = l1 + l2 + l3 - j1 - j3
    return (j1 * (j1 + 1) * (j1 + 2)) // 6 + (j2 * (j2 + 1)) // 2 + j3


def threej000_table_device(N, device=None, stream=None):
    """The N(N+1)(N+2)/6 packed symbols (N > j1 >= j2 >= j3) as a float64
    CUDA tensor, enqueued on `stream` (default: the current stream)."""
    N = int(N)
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    lib = _lib.load()
    with torch.cuda.device(device):
        store = torch.empty(N * (N + 1) * (N + 2) // 6, dtype=torch.float64, device=device)
        s = stream if stream is not None else torch.cuda.current_stream(device)
        _lib.check(lib.dsht_threej000_table(N, store.data_ptr(), s.cuda_stream),
                   "dsht_threej000_table")
    return store


class Wigner3j:
    """Table of (l1 l2 l3; 0 0 0) for l's < Nl (sht/threej000.py:77-116).

    ``table_cache`` (a directory or a ``tablecache.TableCache``): take the
    table from the cache when it holds this Nl, else build it and save it
    there -- the save/load that sht/threej000.py:80-81 describes."""

    def __init__(self, Nl, device=None, table_cache=None):
        self.Nl = int(Nl)
        cache = None
        if table_cache is not None:
            from .tablecache import TableCache
            cache = table_cache if isinstance(table_cache, TableCache) else TableCache(table_cache)
        cached = cache.load_threej(self.Nl) if cache is not None else None
        if cached is not None:
            self._from_store(cached, device)
        else:
            self.store_device = threej000_table_device(self.Nl, device)
            self._store = None
            if cache is not None:
                cache.save_threej(self.Nl, self.store)
        self.from_cache = cached is not None

    def save(self, table_cache):
        """Write the table to a cache directory (or TableCache)."""
        from .tablecache import TableCache
        cache = table_cache if isinstance(table_cache, TableCache) else TableCache(table_cache)
        return cache.save_threej(self.Nl, self.store)

    def _from_store(self, store, device):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.store_device = torch.from_numpy(store).to(device)
        self._store = store

    @classmethod
    def load(cls, Nl, table_cache, device=None):
        """The table from a cache directory; KeyError when it holds no entry."""
        from .tablecache import TableCache
        cache = table_cache if isinstance(table_cache, TableCache) else TableCache(table_cache)
        store = cache.load_threej(int(Nl))
        if store is None:
            raise KeyError(f"no cached 3j table for Nl={Nl} in {cache.root}")
        obj = object.__new__(cls)
        obj.Nl = int(Nl)
        obj._from_store(store, device)
        obj.from_cache = True
        return obj

    @property
    def store(self):
        """The packed table on the host (copied once)."""
        if self._store is None:
            self._store = self.store_device.cpu().numpy()
        return self._store

    def __call__(self, l1, l2, l3):
        """(l1 l2 l3; 0 0 0) for integer l's < Nl."""
        return self.store[self.get_index(l1, l2, l3)]

    def get_index(self, l1, l2, l3):
        return get_index(l1, l2, l3)

    def get_3js(self):
        """The packed table (host numpy array)."""
        return self.store
