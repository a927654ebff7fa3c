"""A continuous-batching transform server on one GPU.

The reference runs one transform per call (sht/sht.py:112-150).  This module
serves a stream of independent transform requests of uneven size and band
limit against resident tables:

  * tables: one resident plan per (Nl, Nx), kept in an LRU; a table is only
    evicted when no request is using it and HBM is needed;
  * slots: up to ``max_inflight`` requests run at once, each on its own HIP
    stream with its own workspace (a clone of the resident plan, see
    DirectSHT.clone), so their kernels overlap on the device;
  * admission: a request is admitted, in arrival order, as soon as a slot is
    free and its HBM footprint (tables if not resident + workspace + input
    and output buffers, from dsht_plan_estimate) fits the budget; requests
    that arrive while others run are admitted the moment one finishes --
    there is no batch boundary to wait for;
  * results: alm copied to pinned host memory on the request's stream; the
    request's latency is submit -> alm on the host.

Each request's result does not depend on what else is in flight: every
transform has its own workspace, the points are sorted in input order within
each segment, and every reduction runs in a fixed order, so a request gives
the same bits alone and in the mix (a serving test checks this).
"""
import collections
import concurrent.futures as cf
import threading
import time

import numpy as np
import torch

from . import _lib
from .sht import DirectSHT


class Request:
    __slots__ = ("rid", "theta", "phi", "wt", "Nl", "Nx", "npoints", "future", "t_submit",
                 "t_admit", "t_done", "need")

    def __init__(self, rid, theta, phi, wt, Nl, Nx):
        self.rid = rid
        self.theta, self.phi, self.wt = theta, phi, wt
        self.Nl, self.Nx = Nl, Nx
        self.npoints = len(theta)
        self.future = cf.Future()
        self.t_submit = time.perf_counter()
        self.t_admit = self.t_done = None
        self.need = 0


class _Slot:
    def __init__(self, device):
        self.stream = torch.cuda.Stream(device=device)
        self.req = None
        self.sht = None
        self.event = None
        self.host_out = None
        self.status = None
        self.keep = None  # device buffers alive until the transform is done


class TransformServer:
    """Serve DirectSHT transforms of uneven (N, Nl) requests, several at once.

    submit(theta, phi, wt, Nl, Nx=None) -> concurrent.futures.Future of the
    healpy-ordered complex128 alm (numpy).  Inputs may be numpy arrays (copied
    to the device on the request's stream) or CUDA float64 tensors (used in
    place).  Nx defaults to nx_factor * Nl.  table_cache: a directory (or
    tablecache.TableCache) whose saved Y tables the server's plans start from
    (tables it builds are saved there).
    """

    def __init__(self, hbm_budget=None, max_inflight=4, dflt_type="float32", xmax=0.75,
                 nx_factor=2, device=None, table_cache=None):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.budget = int(hbm_budget)
        self.dflt_type = dflt_type
        self.xmax = float(xmax)
        self.nx_factor = int(nx_factor)
        self.max_inflight = int(max_inflight)
        self.table_cache = table_cache  # directory of tables saved to disk (tablecache.py)
        self._tables = collections.OrderedDict()   # (Nl, Nx) -> resident DirectSHT (LRU order)
        self._table_bytes = {}                     # (Nl, Nx) -> bytes
        self._idle = collections.defaultdict(list)  # (Nl, Nx) -> idle clones
        self._busy = collections.Counter()          # (Nl, Nx) -> clones in use
        self._reserved = 0                          # bytes held by admitted requests
        self._queue = collections.deque()
        self._slots = [_Slot(self.device) for _ in range(self.max_inflight)]
        self._lock = threading.Lock()
        self._wake = threading.Event()
        self._running = True
        self._next_id = 0
        self.completed = []  # finished Request objects (latency records)
        self._thread = threading.Thread(target=self._loop, name="dsht-server", daemon=True)
        self._thread.start()

    # ---- public ---------------------------------------------------------
    def submit(self, theta, phi, wt, Nl, Nx=None):
        Nl = int(Nl)
        Nx = int(Nx) if Nx is not None else self.nx_factor * Nl
        if not (len(theta) == len(phi) == len(wt)):
            raise AssertionError("theta,phi,wt must be the same length.")
        with self._lock:
            req = Request(self._next_id, theta, phi, wt, Nl, Nx)
            self._next_id += 1
            req.need = self._request_bytes(req)
            if req.need > self.budget:
                req.future.set_exception(MemoryError(
                    f"request needs {req.need} bytes of HBM, budget is {self.budget}"))
                return req.future
            self._queue.append(req)
        self._wake.set()
        return req.future

    def submit_wire(self, frame):
        """Serve one request frame (directsht_amd.wire): returns a Future of
        the response frame (bytes).  Malformed frames, requests over the HBM
        budget and failed transforms give error frames, not exceptions."""
        from . import wire
        out = cf.Future()
        try:
            rid, Nl, Nx, t, p, w = wire.decode_request(frame)
        except ValueError as e:
            out.set_result(wire.encode_error(0, wire.MALFORMED, e))
            return out
        # the frame's arrays are read-only views; the request owns a copy
        fut = self.submit(t.copy(), p.copy(), w.copy(), Nl, Nx if Nx else None)

        def done(f):
            try:
                out.set_result(wire.encode_response(rid, Nl, f.result()))
            except MemoryError as e:
                out.set_result(wire.encode_error(rid, wire.REJECTED, e, Nl))
            except Exception as e:  # the transform's own errors (AssertionError, ...)
                out.set_result(wire.encode_error(rid, wire.FAILED, repr(e), Nl))

        fut.add_done_callback(done)
        return out

    def close(self):
        self._running = False
        self._wake.set()
        self._thread.join()
        for slot in self._slots:
            slot.stream.synchronize()
        self._idle.clear()
        self._tables.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self):
        """Latency percentiles (ms) and aggregate throughput of completed requests."""
        done = [r for r in self.completed if r.t_done is not None]
        if not done:
            return {}
        lat = np.array([(r.t_done - r.t_submit) * 1e3 for r in done])
        t0 = min(r.t_submit for r in done)
        t1 = max(r.t_done for r in done)
        pm = sum(r.npoints * r.Nl * (r.Nl + 1) // 2 for r in done)
        return {"requests": len(done), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                "points_modes_per_s": pm / (t1 - t0), "alm_per_s":
                sum(r.Nl * (r.Nl + 1) // 2 for r in done) / (t1 - t0), "wall_s": t1 - t0}

    # ---- HBM accounting -------------------------------------------------
    def _request_bytes(self, req):
        tb, wb = DirectSHT.estimate(req.Nl, req.Nx, req.npoints, self.dflt_type)
        io = 16 * req.Nl * (req.Nl + 1) // 2
        if not (isinstance(req.theta, torch.Tensor) and req.theta.is_cuda):
            io += 24 * req.npoints
        return wb + io + (0 if (req.Nl, req.Nx) in self._tables else tb)

    @staticmethod
    def _ws_bytes(sht):
        return sht.footprint() - sht._table_bytes_shared

    def _used(self, key=None):
        """HBM in use; the idle workspace that a request for `key` would reuse
        (the largest idle one of that key) is not counted."""
        idle = 0
        for k, cl in self._idle.items():
            sizes = sorted(self._ws_bytes(c) for c in cl)
            if k == key and sizes:
                sizes.pop()
            idle += sum(sizes)
        return sum(self._table_bytes.values()) + self._reserved + idle

    def _make_room(self, need, key):
        """Free idle workspaces, then unused tables (LRU), until need fits."""
        while self._used(key) + need > self.budget:
            victim = next((cl for k, cl in self._idle.items() if cl and (k != key or len(cl) > 1)),
                          None)
            if victim is not None:
                victim.sort(key=self._ws_bytes)
                victim.pop(0)  # DirectSHT.__del__ destroys the clone's workspace
                continue
            k = next((k for k in self._tables if self._busy[k] == 0 and k != key), None)
            if k is None:
                return False
            del self._tables[k]
            del self._table_bytes[k]
            self._idle.pop(k, None)
        return True

    # ---- scheduling -----------------------------------------------------
    def _acquire(self, key):
        if key not in self._tables:
            base = DirectSHT(key[0], key[1], self.xmax, self.dflt_type, device=self.device,
                             table_cache=self.table_cache)
            base._table_bytes_shared = base.footprint()  # no workspace yet: tables only
            self._tables[key] = base
            self._table_bytes[key] = base._table_bytes_shared
        self._tables.move_to_end(key)
        if self._idle[key]:
            self._idle[key].sort(key=self._ws_bytes)
            sht = self._idle[key].pop()  # the largest idle workspace
        else:
            base = self._tables[key]
            sht = base.clone()
            sht._table_bytes_shared = base._table_bytes_shared
        self._busy[key] += 1
        return sht

    def _admit(self):
        while self._queue:
            slot = next((s for s in self._slots if s.req is None), None)
            if slot is None:
                return
            req = self._queue[0]
            key = (req.Nl, req.Nx)
            need = self._request_bytes(req)
            if not self._make_room(need, key):
                if any(s.req is not None for s in self._slots):
                    return  # wait for a running request to free memory
                req.future.set_exception(MemoryError("cannot fit request in the HBM budget"))
                self._queue.popleft()
                continue
            self._queue.popleft()
            req.need = need
            self._reserved += need
            req.t_admit = time.perf_counter()
            try:
                self._launch(slot, req, key)
            except Exception as e:  # device error: fail this request only
                self._reserved -= need
                self._busy[key] -= 1
                slot.req = None
                req.future.set_exception(e)

    def _launch(self, slot, req, key):
        sht = self._acquire(key)
        slot.req, slot.sht = req, sht
        sht.reserve(req.npoints)
        with torch.cuda.device(self.device), torch.cuda.stream(slot.stream):
            keep = []
            args = []
            for a in (req.theta, req.phi, req.wt):
                if isinstance(a, torch.Tensor) and a.is_cuda:
                    args.append(a)
                else:
                    h = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).pin_memory()
                    d = h.to(self.device, non_blocking=True)
                    keep += [h, d]
                    args.append(d)
            alm, status = sht.transform_device(*args, stream=slot.stream)
            host = torch.empty(alm.shape, dtype=alm.dtype, pin_memory=True)
            host.copy_(alm, non_blocking=True)
            st = torch.empty(1, dtype=torch.int32, pin_memory=True)
            st.copy_(status, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        slot.keep = keep + [alm, status]
        slot.host_out, slot.status, slot.event = host, st, ev

    def _finish(self, slot):
        req, sht = slot.req, slot.sht
        req.t_done = time.perf_counter()
        key = (req.Nl, req.Nx)
        st = int(slot.status[0])
        alm = slot.host_out.numpy().copy()
        self._busy[key] -= 1
        self._idle[key].append(sht)
        self._reserved -= req.need
        slot.req = slot.sht = slot.event = slot.host_out = slot.status = slot.keep = None
        self.completed.append(req)
        if st & _lib.DSHT_STATUS_EQU_RANGE:
            req.future.set_exception(AssertionError("theta must be in [ACos[xmax],ACos[-xmax])."))
        else:
            req.future.set_result(alm)  # POLAR_RANGE: zeros, as the reference returns

    def _loop(self):
        while self._running or self._queue or any(s.req is not None for s in self._slots):
            with self._lock:
                progressed = False
                for slot in self._slots:
                    if slot.req is not None and slot.event.query():
                        self._finish(slot)
                        progressed = True
                self._admit()
                busy = any(s.req is not None for s in self._slots)
            if not progressed:
                if busy:
                    time.sleep(5e-5)
                else:
                    self._wake.wait(0.01)
                    self._wake.clear()
