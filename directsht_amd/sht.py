"""DirectSHT: drop-in for the reference's sht/sht.py:30-284 on MI355X.

``DirectSHT(Nell, Nx, xmax=0.75, dflt_type='float64')(thetas, phis, weights)``
returns healpy-ordered complex alm, a_lm = sum_i w_i Y*_lm(theta_i, phi_i) with
the reference's cubic-Hermite spline approximation of Y_lm in cos(theta) and
its 90-degree rotation of polar points.  All compute runs in hand-written HIP
kernels (libdsht.so, include/dsht.h); torch only provides device memory and
the current stream.

Differences from the reference, all documented:
  * ``dflt_type`` selects the precision of the table contraction ('float64':
    FP64 tables; 'float32': FP32 tables with FP64 accumulation).  The tables
    are always built in FP64 on the FP64 grid (the reference's numpy path
    builds FP64 tables on a grid of dtype dflt_type).
  * ``null_unphysical`` is accepted and ignored, as in the reference's numpy
    path (sht/legendre_py.py:63-65): the healpy-row tables have no l < m rows.
  * Timing prints of basic_sht (sht/sht.py:224-282) are off unless
    ``verbose=True`` is passed to basic_sht.
  * Inputs may be numpy arrays (result: numpy complex128, as the reference) or
    torch tensors on the GPU (result: a complex128 torch tensor on the same
    device, no host round trip).
"""
import ctypes as ct
import time

import numpy as np
import torch

from . import _lib


def _as_device(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)


class DirectSHT:
    """Brute-force spherical harmonic transforms of point sets (sht/sht.py:30)."""

    def __init__(self, Nell, Nx, xmax=0.75, dflt_type="float64", null_unphysical=True,
                 device=None, table_cache=None):
        """Build the FP64 spline tables on the GPU (sht/sht.py:32-56).

        ``table_cache`` (a directory or a ``tablecache.TableCache``; no
        counterpart in the reference): take the Y table from the cache when it
        holds this (Nl, Nx, xmax), else build it and save it there."""
        if dflt_type not in ("float64", "float32"):
            raise ValueError("dflt_type must be 'float64' or 'float32'")
        self.Nell, self.Nx, self.xmax = int(Nell), int(Nx), float(xmax)
        self.Nlm = (self.Nell * (self.Nell + 1)) // 2
        self.dflt_type = dflt_type
        self.null_unphysical = null_unphysical
        xx = np.arange(self.Nx, dtype="float64")
        xx *= self.xmax / float(self.Nx - 1)
        self.x = xx
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        self._precision = _lib.DSHT_FP64 if dflt_type == "float64" else _lib.DSHT_FP32
        lib = _lib.load()
        plan = ct.c_void_p()
        cache = None
        if table_cache is not None:
            from .tablecache import TableCache
            cache = table_cache if isinstance(table_cache, TableCache) else TableCache(table_cache)
        cached = cache.load_ylm(self.Nell, self.Nx, self.xmax) if cache is not None else None
        with torch.cuda.device(self.device):
            if cached is not None:
                Yv = np.ascontiguousarray(cached)
                _lib.check(lib.dsht_plan_create_from_tables(
                    self.Nell, self.Nx, self.xmax, self._precision, Yv.ctypes.data,
                    ct.byref(plan)), "dsht_plan_create_from_tables")
            else:
                _lib.check(lib.dsht_plan_create(self.Nell, self.Nx, self.xmax, self._precision,
                                                ct.byref(plan)), "dsht_plan_create")
        self._plan = plan
        self._status = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._Yv = self._Yd = None
        self.from_cache = cached is not None
        if cache is not None and cached is None:
            cache.save_ylm(self.Nell, self.Nx, self.xmax, self.Yv)

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.load().dsht_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None

    # ---- concurrency helpers (no counterpart in the reference) ------------
    def clone(self):
        """A DirectSHT over the same resident tables with its own workspace:
        transforms of the clone and of the original may run concurrently on
        different streams.  The tables are freed with the last sharer."""
        other = object.__new__(DirectSHT)
        other.__dict__.update({k: v for k, v in self.__dict__.items()
                               if k not in ("_plan", "_status", "_Yv", "_Yd")})
        plan = ct.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().dsht_plan_clone(self._plan, ct.byref(plan)), "dsht_plan_clone")
        other._plan = plan
        other._status = torch.zeros(1, dtype=torch.int32, device=self.device)
        other._Yv = other._Yd = None
        return other

    def reserve(self, npoints, nsets=1, point_sets=True, weight_sets=True):
        """Size the workspace for transforms of up to ``npoints`` points (per
        set) and ``nsets`` sets now, so later transforms neither allocate nor
        synchronize: point-set transforms (``__call__``, ``transform_device``,
        ``transform_many_device``) and/or weight-set transforms
        (``transform_sets_device``)."""
        kinds = (1 if point_sets else 0) | (2 if weight_sets else 0)
        if not kinds:
            raise ValueError("reserve: nothing to reserve for")
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().dsht_plan_reserve_for(self._plan, int(npoints), int(nsets),
                                                         kinds), "dsht_plan_reserve")

    def footprint(self):
        """HBM bytes held: tables (shared with clones) + own workspace."""
        b = ct.c_int64()
        _lib.check(_lib.load().dsht_plan_footprint(self._plan, ct.byref(b)), "dsht_plan_footprint")
        return b.value

    @staticmethod
    def estimate(Nell, Nx, npoints, dflt_type="float32"):
        """(table_bytes, workspace_bytes) a plan and a transform of
        ``npoints`` points need -- what a server admits against."""
        prec = _lib.DSHT_FP64 if dflt_type == "float64" else _lib.DSHT_FP32
        tb, wb = ct.c_int64(), ct.c_int64()
        _lib.check(_lib.load().dsht_plan_estimate(int(Nell), int(Nx), prec, int(npoints),
                                                  ct.byref(tb), ct.byref(wb)), "dsht_plan_estimate")
        return tb.value, wb.value

    # ---- table access (sht/sht.py:56,69-93) -----------------------------
    def _fetch_tables(self):
        Yv = np.empty((self.Nlm, self.Nx), dtype=np.float64)
        Yd = np.empty((self.Nlm, self.Nx), dtype=np.float64)
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().dsht_plan_get_tables(self._plan, Yv.ctypes.data, Yd.ctypes.data),
                       "dsht_plan_get_tables")
        self._Yv, self._Yd = Yv, Yd

    @property
    def Yv(self):
        """Y_lm(x_j, 0), shape (Nlm, Nx), healpy row order (copied from the GPU)."""
        if self._Yv is None:
            self._fetch_tables()
        return self._Yv

    @property
    def Yd(self):
        """dY_lm/dx at x_j, shape (Nlm, Nx), healpy row order."""
        if self._Yd is None:
            self._fetch_tables()
        return self._Yd

    def indx(self, ell, m):
        """healpy index of (ell, m) -- sht/sht.py:58-67."""
        return (m * (2 * self.Nell - 1 - m)) // 2 + ell

    def get_Ylm(self, ell, m):
        """Y_lm(x_i, 0) on the grid -- sht/sht.py:69-80 (numpy-path layout)."""
        return self.Yv[self.indx(ell, m), :]

    def get_dYlm(self, ell, m):
        """dY_lm/dx on the grid -- sht/sht.py:82-93."""
        return self.Yd[self.indx(ell, m), :]

    def alm2cl(self, alm):
        """[(Re a_l0)^2 + 2 sum_{m>0} |a_lm|^2] / (2l+1) -- sht/sht.py:95-109.

        Computed on the device (dsht_alm2cl).  alm: Nlm complex values
        (numpy -> numpy result, CUDA tensor -> CUDA tensor), or (S, Nlm) for
        the S auto-spectra (S, Nl)."""
        return self._spectra(alm, cross=False)

    def cross_spectra(self, alms):
        """Cross-spectra of S alm arrays (S, Nlm): (S(S+1)/2, Nl) for the pairs
        (s, t), s <= t, in row order; C^{st}_l = [Re a^s_l0 Re a^t_l0 +
        2 sum_{m>0} Re(a^s_lm conj(a^t_lm))] / (2l+1), healpy's
        alm2cl(alm1, alm2) with the m = 0 convention of alm2cl.  On the device."""
        return self._spectra(alms, cross=True)

    def _spectra(self, alm, cross):
        on_device = isinstance(alm, torch.Tensor) and alm.is_cuda
        a = alm if isinstance(alm, torch.Tensor) else torch.from_numpy(np.asarray(alm))
        one = a.dim() == 1
        a = a.reshape(-1, a.shape[-1]).to(device=self.device, dtype=torch.complex128).contiguous()
        S, n = a.shape
        if n != self.Nlm:
            raise ValueError(f"alm has {n} values per set, expected Nlm={self.Nlm}")
        if cross and S > _lib.MAX_SETS:
            raise ValueError(f"cross_spectra takes at most {_lib.MAX_SETS} sets")
        lib = _lib.load()
        out = []
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device)
            for s0 in range(0, S, _lib.MAX_SETS):
                grp = a[s0:s0 + _lib.MAX_SETS]
                k = grp.shape[0]
                npairs = k * (k + 1) // 2 if cross else k
                cl = torch.empty((npairs, self.Nell), dtype=torch.float64, device=self.device)
                _lib.check(lib.dsht_alm2cl(self.Nell, k, grp.data_ptr(), self.Nlm, cl.data_ptr(),
                                           int(cross), stream.cuda_stream), "dsht_alm2cl")
                out.append(cl)
        cl = torch.cat(out) if len(out) > 1 else out[0]
        if one and not cross:
            cl = cl[0]
        return cl if on_device else cl.cpu().numpy()

    # ---- tuning -----------------------------------------------------------
    def configure(self, scatter=None, chunk=None, contract=None):
        """Select the scatter engine of FP32 plans ('auto', the default: per
        transform, 'spread' where Nl <= 2048 and the points average at least
        64 (single transforms on grids of 512 or more node pairs, Nl <= 1024),
        256 (other transforms) or 640 (weight-set batches) per segment of
        the grid, else 'matrix'; 'matrix': fp16x3 MFMA direct sums,
        'vector': FP32 VALU direct sums, 'spread': non-uniform FFTs,
        ~1e-7 relative), the sorted points per scatter work item (the
        largest; small transforms use smaller items), and the table
        contraction of batched FP32 transforms ('matrix': fp16x3 MFMA GEMM
        over all sets, 'vector': FP32 GEMV per set).  Results do not depend
        on these beyond rounding (and the spreading kernel's ~1e-7)."""
        lib = _lib.load()
        if contract is not None:
            v = {"matrix": _lib.DSHT_CONTRACT_MATRIX, "vector": _lib.DSHT_CONTRACT_VECTOR}[contract]
            _lib.check(lib.dsht_plan_configure(self._plan, _lib.DSHT_OPT_CONTRACT, v), "configure")
        if scatter is not None:
            v = {"matrix": _lib.DSHT_SCATTER_MATRIX, "vector": _lib.DSHT_SCATTER_VECTOR,
                 "spread": _lib.DSHT_SCATTER_SPREAD, "auto": _lib.DSHT_SCATTER_AUTO}[scatter]
            _lib.check(lib.dsht_plan_configure(self._plan, _lib.DSHT_OPT_SCATTER, v), "configure")
        if chunk is not None:
            _lib.check(lib.dsht_plan_configure(self._plan, _lib.DSHT_OPT_CHUNK, int(chunk)), "configure")

    def scatter_engine(self, npoints):
        """The scatter engine ('matrix', 'vector' or 'spread') a transform of
        npoints points takes on this plan."""
        e = ct.c_int()
        _lib.check(_lib.load().dsht_plan_engine(self._plan, int(npoints), ct.byref(e)),
                   "dsht_plan_engine")
        return {_lib.DSHT_SCATTER_MATRIX: "matrix", _lib.DSHT_SCATTER_VECTOR: "vector",
                _lib.DSHT_SCATTER_SPREAD: "spread"}[e.value]

    # ---- stage timing (HIP events on the transform's stream) -------------
    def set_timing(self, enable=True):
        """Start (and reset) or stop per-stage event timing of transforms."""
        _lib.check(_lib.load().dsht_plan_set_timing(self._plan, int(bool(enable))),
                   "dsht_plan_set_timing")

    def get_timing(self):
        """(dict stage -> summed ms, number of transforms) since set_timing."""
        ms = (ct.c_double * len(_lib.STAGES))()
        n = ct.c_int64()
        _lib.check(_lib.load().dsht_plan_get_timing(self._plan, ms, ct.byref(n)),
                   "dsht_plan_get_timing")
        return dict(zip(_lib.STAGES, list(ms))), n.value

    def debug_node_coefficients(self):
        """Node coefficients of the last single-set transform, shape
        (Nl, 2, Npair, 24) with Npair = 2 ceil(Nx/4) (nodes past Nx are padding):
        [m][parity][node pair][kind*8 + q*2 + node],
        kinds c, f, g (see dsht_plan_debug_rhs in include/dsht.h)."""
        npair = 2 * ((self.Nx + 3) // 4)
        dt = np.float32 if self._precision == _lib.DSHT_FP32 else np.float64
        # the device layout [pair][branch][m][P][kind][re|im][node]
        raw = np.empty((npair, 2, self.Nell, 2, 3, 2, 2), dtype=dt)
        _lib.check(_lib.load().dsht_plan_debug_rhs(self._plan, raw.ctypes.data, raw.nbytes),
                   "dsht_plan_debug_rhs")
        # -> [m][P][pair][kind][branch][re|im][node]
        return np.ascontiguousarray(raw.transpose(2, 3, 0, 4, 1, 5, 6)).reshape(
            self.Nell, 2, npair, 24)

    # ---- transforms -----------------------------------------------------
    def transform_device(self, theta, phi, wt, out=None, mode=_lib.DSHT_MODE_CALL, stream=None):
        """Enqueue one transform of device-resident float64 tensors.

        Returns (alm, status) where alm is a complex128 tensor of length Nlm
        (``out`` if given) and status a device int32 tensor of DSHT_STATUS_*
        bits.  Nothing is synchronized.
        """
        n = theta.numel()
        if out is None:
            out = torch.empty(self.Nlm, dtype=torch.complex128, device=self.device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().dsht_transform(
                self._plan, n, theta.data_ptr(), phi.data_ptr(), wt.data_ptr(), out.data_ptr(),
                self._status.data_ptr(), mode, stream.cuda_stream), "dsht_transform")
        return out, self._status

    def transform_sets_device(self, theta, phi, W, out=None, mode=_lib.DSHT_MODE_CALL, stream=None,
                              status=None):
        """Enqueue transforms of many weight sets on shared points: W is a
        (S, N) float64 CUDA tensor; returns (alm (S, Nlm) complex128, status
        (S,) int32), nothing synchronized.  The points are binned once and the
        tables read once per group of up to 64 sets."""
        n = theta.numel()
        S = W.shape[0]
        W = W.contiguous()
        if out is None:
            out = torch.empty((S, self.Nlm), dtype=torch.complex128, device=self.device)
        if status is None:
            status = torch.zeros(S, dtype=torch.int32, device=self.device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        lib = _lib.load()
        with torch.cuda.device(self.device):
            for s0 in range(0, S, _lib.MAX_SETS):
                k = min(_lib.MAX_SETS, S - s0)
                _lib.check(lib.dsht_transform_sets(
                    self._plan, n, theta.data_ptr(), phi.data_ptr(), k, W[s0].data_ptr(), W.stride(0),
                    out[s0].data_ptr(), status[s0:].data_ptr(), mode, stream.cuda_stream),
                    "dsht_transform_sets")
        return out, status

    def transform_many_device(self, sets, out=None, mode=_lib.DSHT_MODE_CALL, stream=None,
                              status=None):
        """Enqueue transforms of many point sets at this (Nell, Nx): sets is a
        list of (theta, phi, w) float64 CUDA tensors; returns (alm (S, Nlm),
        status (S,)), nothing synchronized.  The table contraction and the
        polar rotation run once per group of up to 64 sets."""
        S = len(sets)
        if out is None:
            out = torch.empty((S, self.Nlm), dtype=torch.complex128, device=self.device)
        if status is None:
            status = torch.zeros(S, dtype=torch.int32, device=self.device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        lib = _lib.load()
        with torch.cuda.device(self.device):
            for s0 in range(0, S, _lib.MAX_SETS):
                grp = sets[s0:s0 + _lib.MAX_SETS]
                k = len(grp)
                Np = (ct.c_int64 * k)(*[g[0].numel() for g in grp])
                th = (ct.c_void_p * k)(*[g[0].data_ptr() for g in grp])
                ph = (ct.c_void_p * k)(*[g[1].data_ptr() for g in grp])
                ww = (ct.c_void_p * k)(*[g[2].data_ptr() for g in grp])
                _lib.check(lib.dsht_transform_many(self._plan, k, Np, th, ph, ww, out[s0].data_ptr(),
                                                   status[s0:].data_ptr(), mode, stream.cuda_stream),
                           "dsht_transform_many")
        return out, status

    def _run(self, t, p, w, mode):
        on_device = isinstance(t, torch.Tensor) and t.is_cuda
        n = len(t)
        assert n == len(p) and len(p) == len(w), "theta,phi,wt must be the same length."
        td, pd, wd = (_as_device(a, self.device) for a in (t, p, w))
        alm, status = self.transform_device(td, pd, wd, mode=mode)
        st = int(status.item())  # synchronizes the stream
        return (alm if on_device else alm.cpu().numpy()), st

    def __call__(self, t, p, w):
        """alm of points (t, p) with weights w -- sht/sht.py:112-150.

        w may also be 2-D, shape (S, N): S weight sets on the same points,
        returning alm of shape (S, Nlm) (one binning, one table read)."""
        if getattr(w, "ndim", 1) == 2:
            return self._call_sets(t, p, w)
        alm, st = self._run(t, p, w, _lib.DSHT_MODE_CALL)
        if st & _lib.DSHT_STATUS_POLAR_RANGE:
            # sht/sht.py:141-147: caught, printed, zeros returned
            print("AssertionError in basic_sht: theta must be in [ACos[xmax],ACos[-xmax]).")
            if self.xmax < np.sin(np.pi / 4):
                print("Warning: xmax is smaller than sin(pi/4), so some points may"
                      " still lie in the polar cap after rotation by 90 degrees."
                      " Trying increasing it.")
            return alm
        if st & _lib.DSHT_STATUS_EQU_RANGE:
            raise AssertionError("theta must be in [ACos[xmax],ACos[-xmax]).")
        return alm

    def _call_sets(self, t, p, W):
        on_device = isinstance(t, torch.Tensor) and t.is_cuda
        n = len(t)
        assert n == len(p) and W.shape[1] == n, "theta,phi,wt must be the same length."
        td, pd = (_as_device(a, self.device) for a in (t, p))
        Wd = W.to(self.device, torch.float64) if isinstance(W, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(W, dtype=np.float64)).to(self.device)
        alm, status = self.transform_sets_device(td, pd, Wd)
        st = int(status.max().item()) if status.numel() else 0
        if st & _lib.DSHT_STATUS_EQU_RANGE:
            raise AssertionError("theta must be in [ACos[xmax],ACos[-xmax]).")
        if st & _lib.DSHT_STATUS_POLAR_RANGE:
            print("AssertionError in basic_sht: theta must be in [ACos[xmax],ACos[-xmax]).")
        return alm if on_device else alm.cpu().numpy()

    def basic_sht(self, theta, phi, wt, reg_factor=1.0, verbose=False):
        """alm of points that all satisfy |cos theta| < xmax -- sht/sht.py:153-283."""
        if len(theta) == 0 and len(phi) == 0 and len(wt) == 0:
            raise ValueError("The theta array seems to be empty!")
        t0 = time.time()
        w = wt * reg_factor
        alm, st = self._run(theta, phi, w, _lib.DSHT_MODE_BASIC)
        if st & _lib.DSHT_STATUS_EQU_RANGE:
            raise AssertionError("theta must be in [ACos[xmax],ACos[-xmax]).")
        if verbose:
            print("Direct SHT took ", time.time() - t0, " seconds.", flush=True)
        return alm / reg_factor
