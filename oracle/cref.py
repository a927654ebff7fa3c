"""ctypes drivers for the reference's own C code, oracle/_ref/sht_helper.so
(and, at the end, oracle/_ref/threej000_helper.so).

TEST INFRASTRUCTURE ONLY.  The shared object is compiled by oracle/Makefile
from csht/sht_helper.c in the reference checkout (it is never copied into this
repository).  The driver mirrors the reference's own ctypes wrapper
csht/sht.py:22-153: make_table fills caller-owned healpy-row tables
(csht/sht_helper.c:18-116); do_transform (csht/sht_helper.c:176-286) takes one
sign of cos(theta), sorted by |cos(theta)|, and returns Sigma Y cos / Sigma Y sin
from which alm = carr - 1j*sarr; the caller applies (-1)^(l-m) for cos < 0
(csht/sht.py:101-121).

The contraction loop of do_transform (csht/sht_helper.c:258-262) shares the
thread index between OpenMP threads, a data race; results are only reliable
with one thread, which is why golden fixtures are written with
OMP_NUM_THREADS=1.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(_HERE, "_ref", "sht_helper.so")


def available():
    return os.path.exists(SO_PATH)


def _lib():
    lib = ct.CDLL(SO_PATH)
    dp = ct.POINTER(ct.c_double)
    lib.make_table.argtypes = [ct.c_int, ct.c_int, ct.c_double, dp, dp]
    lib.make_table.restype = ct.c_int
    lib.do_transform.argtypes = [ct.c_int, ct.c_int, ct.c_double, dp, dp, ct.c_int,
                                 dp, dp, dp, dp, dp]
    lib.do_transform.restype = ct.c_int
    return lib


def _p(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class CRef:
    """The csht DirectSHT (csht/sht.py:22-153) on the compiled reference C."""

    def __init__(self, Nell, Nx, xmax=0.75):
        self.Nell, self.Nx, self.xmax = Nell, Nx, xmax
        self.Nlm = Nell * (Nell + 1) // 2
        self.lib = _lib()
        self.Yv = np.zeros(self.Nlm * Nx)
        self.Yd = np.zeros(self.Nlm * Nx)
        rc = self.lib.make_table(Nell, Nx, xmax, _p(self.Yv), _p(self.Yd))
        if rc != 0:
            raise RuntimeError("make_table failed")

    def tables(self):
        return self.Yv.reshape(self.Nlm, self.Nx), self.Yd.reshape(self.Nlm, self.Nx)

    def _one_sign(self, cost, phi, wt):
        carr = np.zeros(self.Nlm)
        sarr = np.zeros(self.Nlm)
        cost, phi, wt = (np.ascontiguousarray(a, dtype=np.float64) for a in (cost, phi, wt))
        rc = self.lib.do_transform(self.Nell, self.Nx, self.xmax, _p(self.Yv), _p(self.Yd),
                                   len(cost), _p(cost), _p(phi), _p(wt), _p(carr), _p(sarr))
        if rc != 0:
            raise RuntimeError("do_transform failed")
        return carr - 1j * sarr

    def __call__(self, theta, phi, wt):
        """csht/sht.py:78-140 (equatorial points only: |cos theta| < xmax)."""
        alm = np.zeros(self.Nlm, dtype=np.complex128)
        cost = np.cos(theta)
        neg = np.nonzero(cost < 0)[0]
        if len(neg) > 0:
            ii = neg[np.argsort(-cost[neg])]
            a = self._one_sign(cost[ii], phi[ii], wt[ii])
            ell = np.concatenate([np.arange(m, self.Nell) for m in range(self.Nell)])
            m = np.concatenate([np.full(self.Nell - m, m) for m in range(self.Nell)])
            alm += a * (-1.0) ** (ell - m)
        pos = np.nonzero(cost >= 0)[0]
        if len(pos) > 0:
            ii = pos[np.argsort(cost[pos])]
            alm += self._one_sign(cost[ii], phi[ii], wt[ii])
        return alm


TJ_PATH = os.path.join(_HERE, "_ref", "threej000_helper.so")


def threej_available():
    return os.path.exists(TJ_PATH)


def threej000_table(N):
    """The reference's C Wigner-3j table, make_table(Nl, store)
    (csht/threej000_helper.c:62-77), compiled by oracle/Makefile:
    N(N+1)(N+2)/6 values (j1 j2 j3; 0 0 0), N > j1 >= j2 >= j3, at
    j1(j1+1)(j1+2)/6 + j2(j2+1)/2 + j3."""
    lib = ct.CDLL(TJ_PATH)
    lib.make_table.argtypes = [ct.c_int, ct.POINTER(ct.c_double)]
    lib.make_table.restype = ct.c_int
    store = np.zeros(N * (N + 1) * (N + 2) // 6)
    if lib.make_table(N, _p(store)) != 0:
        raise RuntimeError("make_table failed")
    return store
