"""On-disk cache of the spline-node tables and of the Wigner-3j table.

The reference recomputes its tables per instance (sht/sht.py:50-56); its
``Wigner3j`` docstring promises save/load (sht/threej000.py:80-81) without
implementing it.  Here a cache directory holds

* the FP64 Y table of a DirectSHT plan in healpy row order [Nlm][Nx] (the
  reference's ``Yv`` attribute), keyed by (Nl, Nx, xmax, dtype); one entry
  serves both plan precisions, since an FP32 plan's table is the FP64 one
  rounded (dsht_plan_create_from_tables);
* the packed (l1 l2 l3; 0 0 0) table of ``Wigner3j(N)``, keyed by N.

Each entry is a ``.npy`` array (loaded with ``allow_pickle=False``) beside a
JSON header naming the format, the key, the shape and an xxh3-64 checksum of
the array's bytes; both are written to temporary names and renamed into
place, so a reader sees a whole entry or none.  ``load_*`` returns None when
the entry is missing and raises ``ValueError`` when it exists but does not
match its key or checksum.

The device builds these tables in milliseconds (Nl=1000, Nx=2048: 6 ms for
the FP32 table; Wigner3j(1127): 0.4 ms), while the FP64 Y table at that size
is 8.2 GB on disk: the cache is for callers that keep tables beside results
or run without the build (it is off unless a directory is given).
"""
import json
import os
import tempfile

import numpy as np
import xxhash

FORMAT = "dsht-table-cache-1"


def _digest(a):
    return xxhash.xxh3_64_hexdigest(memoryview(np.ascontiguousarray(a)).cast("B"))


class TableCache:
    """A directory of cached tables."""

    def __init__(self, root):
        self.root = os.fspath(root)
        os.makedirs(self.root, exist_ok=True)

    # ---- keys -------------------------------------------------------------
    @staticmethod
    def ylm_key(Nl, Nx, xmax):
        return {"kind": "ylm", "Nl": int(Nl), "Nx": int(Nx), "xmax": float(xmax).hex(),
                "dtype": "float64"}

    @staticmethod
    def threej_key(N):
        return {"kind": "threej000", "N": int(N), "dtype": "float64"}

    def _stem(self, key):
        if key["kind"] == "ylm":
            name = f"ylm_Nl{key['Nl']}_Nx{key['Nx']}_x{key['xmax']}"
        else:
            name = f"threej000_N{key['N']}"
        return os.path.join(self.root, name.replace("+", "p").replace("-", "m"))

    # ---- entries ----------------------------------------------------------
    def _save(self, key, array):
        a = np.ascontiguousarray(array, dtype=np.float64)
        stem = self._stem(key)
        head = {"format": FORMAT, "key": key, "shape": list(a.shape), "xxh3_64": _digest(a)}
        fd, tmp_npy = tempfile.mkstemp(dir=self.root, suffix=".npy.tmp")
        with os.fdopen(fd, "wb") as f:
            np.save(f, a, allow_pickle=False)
        fd, tmp_json = tempfile.mkstemp(dir=self.root, suffix=".json.tmp")
        with os.fdopen(fd, "w") as f:
            json.dump(head, f)
        os.replace(tmp_npy, stem + ".npy")
        os.replace(tmp_json, stem + ".json")
        return stem

    def _load(self, key, mmap=False):
        stem = self._stem(key)
        if not (os.path.exists(stem + ".json") and os.path.exists(stem + ".npy")):
            return None
        with open(stem + ".json") as f:
            head = json.load(f)
        if head.get("format") != FORMAT or head.get("key") != key:
            raise ValueError(f"table cache entry {stem}: header does not match the key {key}")
        a = np.load(stem + ".npy", mmap_mode="r" if mmap else None, allow_pickle=False)
        if list(a.shape) != head["shape"] or a.dtype != np.float64:
            raise ValueError(f"table cache entry {stem}: shape or dtype differs from its header")
        if _digest(a) != head["xxh3_64"]:
            raise ValueError(f"table cache entry {stem}: checksum mismatch")
        return a

    def save_ylm(self, Nl, Nx, xmax, Yv):
        Nlm = int(Nl) * (int(Nl) + 1) // 2
        if np.shape(Yv) != (Nlm, int(Nx)):
            raise ValueError(f"Yv must have shape ({Nlm}, {Nx})")
        return self._save(self.ylm_key(Nl, Nx, xmax), Yv)

    def load_ylm(self, Nl, Nx, xmax):
        """The cached FP64 Y table [Nlm][Nx] or None."""
        return self._load(self.ylm_key(Nl, Nx, xmax), mmap=True)

    def save_threej(self, N, store):
        n = int(N)
        if np.shape(store) != (n * (n + 1) * (n + 2) // 6,):
            raise ValueError("the packed 3j table must hold N(N+1)(N+2)/6 values")
        return self._save(self.threej_key(N), store)

    def load_threej(self, N):
        """The cached packed 3j table or None."""
        return self._load(self.threej_key(N))
