"""Many independent transforms of uneven size and band limit in one launch.

A batch is a list of point sets (device-resident theta, phi, w; each with its
own Nl, Nx).  BatchTransform prepares everything once -- resident tables per
(Nl, Nx), a tuned configuration per tables (the one tuned for the size bucket
of their largest set), workspaces reserved to size, output buffers -- and
captures the whole batch as one HIP graph.  Sets that share tables go through
one dsht_transform_many call (up to 64 sets): each set is binned, scattered and folded
into its own slot (sparse sets in groups on the matrix engine, dense ones
alone on the spreading scatter when the call's engine is 'auto'), and the table contraction (a matrix-core GEMM over all
slots) and the polar rotation run once for the call, so the tables are
streamed from HBM once per call instead of once per set.  Calls are spread over several streams
(longest first, onto the least loaded stream) so they overlap, and the graph
is replayed with a single launch.  Nothing in a replay allocates or
synchronizes.

Autotuner: for each (Nl, Nx, size bucket) it is asked about it times every
candidate (scatter engine, points per scatter work item) on a synthetic
uniform-sphere set of the bucket's size and keeps the fastest; results are
cached per process (and optionally in a JSON file).
"""
import json
import math
import os

import torch

from . import _lib
from .sht import DirectSHT

# (scatter engine, points per work item; None: the engine's default).  On
# 'auto' a dsht_transform_many call sends its dense sets one by one through
# the spreading scatter and groups the others on the matrix engine.
CANDIDATES = (("matrix", 1024), ("matrix", 2048), ("matrix", 4096), ("vector", 2048),
              ("auto", None))


def size_bucket(npoints):
    """Half-decade bucket of a point count (1e4..3.2e4 -> 8, 3.2e4..1e5 -> 9, ...)."""
    return int(math.floor(2.0 * math.log10(max(int(npoints), 1))))


class Autotuner:
    """Pick (scatter engine, chunk) per (Nl, Nx, size bucket, precision) by timing."""

    def __init__(self, cache_file=None, reps=3, candidates=CANDIDATES):
        self.cache_file = cache_file
        self.reps = int(reps)
        self.candidates = tuple(candidates)
        self.table = {}
        if cache_file and os.path.exists(cache_file):
            with open(cache_file) as f:
                for k, v in json.load(f).items():
                    self.table[k] = tuple(v)

    @staticmethod
    def key(Nl, Nx, bucket, dflt_type):
        return f"{Nl}:{Nx}:{bucket}:{dflt_type}"

    def choose(self, base, npoints):
        """Best (engine, chunk) for transforms of about npoints points with the
        tables of `base` (a DirectSHT); timed on first use."""
        b = size_bucket(npoints)
        k = self.key(base.Nell, base.Nx, b, base.dflt_type)
        if k in self.table:
            return self.table[k]
        n = int(round(10 ** ((b + 0.5) / 2.0)))  # geometric middle of the bucket
        g = torch.Generator(device=base.device).manual_seed(1234 + b)
        th = torch.acos(2 * torch.rand(n, generator=g, device=base.device, dtype=torch.float64) - 1)
        ph = 2 * math.pi * torch.rand(n, generator=g, device=base.device, dtype=torch.float64)
        w = torch.ones(n, device=base.device, dtype=torch.float64)
        out = torch.empty(base.Nlm, dtype=torch.complex128, device=base.device)
        best, best_t = None, float("inf")
        for engine, chunk in self.candidates:
            if engine == "vector" and base.dflt_type != "float32":
                continue
            sht = base.clone()
            sht.configure(scatter=engine, chunk=chunk)  # chunk None: the default
            sht.reserve(n)
            sht.transform_device(th, ph, w, out=out)  # warm-up
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(self.reps):
                sht.transform_device(th, ph, w, out=out)
            e1.record()
            e1.synchronize()
            t = e0.elapsed_time(e1) / self.reps
            if t < best_t:
                best, best_t = (engine, chunk), t
            del sht
        self.table[k] = best
        if self.cache_file:
            with open(self.cache_file, "w") as f:
                json.dump(self.table, f)
        return best


class BatchTransform:
    """Capture transforms of many point sets into one HIP graph.

    sets: list of (theta, phi, wt, Nl) or (theta, phi, wt, Nl, Nx) with CUDA
    float64 tensors (kept referenced; their contents may change between
    replays).  run() replays the graph on the current stream and returns the
    list of complex128 alm tensors (overwritten by every replay).
    """

    def __init__(self, sets, dflt_type="float32", xmax=0.75, nstreams=4, tuner=None,
                 nx_factor=2, device=None, graph=True, split_calls=False):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        self.dflt_type = dflt_type
        self.tuner = tuner if tuner is not None else Autotuner()
        self.sets = []
        for s in sets:
            th, ph, w, Nl = s[:4]
            Nx = s[4] if len(s) > 4 else nx_factor * int(Nl)
            self.sets.append((th, ph, w, int(Nl), int(Nx)))
        # resident tables
        self.bases = {}
        for (_, _, _, Nl, Nx) in self.sets:
            if (Nl, Nx) not in self.bases:
                self.bases[(Nl, Nx)] = DirectSHT(Nl, Nx, xmax, dflt_type, device=self.device)
        # tuned configuration per (tables): the one tuned for the largest set
        # of those tables (the scatter of the large sets dominates), so that
        # all sets sharing tables can share a call and its table stream
        by_tables = {}
        for i, s in enumerate(self.sets):
            by_tables.setdefault((s[3], s[4]), []).append(i)
        self.config = [None] * len(self.sets)
        for tk, idx in by_tables.items():
            big = max(idx, key=lambda i: self.sets[i][0].numel())
            cfg = tuple(self.tuner.choose(self.bases[tk], self.sets[big][0].numel()))
            for i in idx:
                self.config[i] = cfg
        # calls: sets with the same tables, up to MAX_SETS per
        # dsht_transform_many call (one table stream for the whole call).
        # With split_calls, a call much longer than a stream's fair share is
        # split in two (alternating sets by size), trading another table
        # stream for balance across the streams (off by default: on the
        # uneven 200-set batch it was slower)
        self.calls = []
        for tk, idx in by_tables.items():
            idx.sort(key=lambda i: -self.sets[i][0].numel())
            key = tk + self.config[idx[0]]
            for c0 in range(0, len(idx), _lib.MAX_SETS):
                self.calls.append((key, idx[c0:c0 + _lib.MAX_SETS]))
        nstr = max(1, int(nstreams))
        while split_calls and nstr > 1 and len(self.calls) < 4 * nstr:
            costs = [self._call_cost(c) for c in self.calls]
            j = max(range(len(self.calls)), key=lambda q: costs[q])
            key, idx = self.calls[j]
            if len(idx) < 2 or costs[j] <= 1.25 * sum(costs) / nstr:
                break
            self.calls[j] = (key, idx[0::2])
            self.calls.append((key, idx[1::2]))
        # assignment of calls to streams: longest first onto the least loaded
        cost = [self._call_cost(c) for c in self.calls]
        order = sorted(range(len(self.calls)), key=lambda i: -cost[i])
        self.nstreams = max(1, min(int(nstreams), len(self.calls)))
        load = [0.0] * self.nstreams
        self.assign = [[] for _ in range(self.nstreams)]
        for i in order:
            j = min(range(self.nstreams), key=lambda q: load[q])
            self.assign[j].append(i)
            load[j] += cost[i]
        # one workspace per (stream, tables, configuration), reserved for its
        # largest set and call
        self.ws = {}
        for j, calls in enumerate(self.assign):
            for ci in calls:
                key, idx = self.calls[ci]
                wk = (j,) + key
                if wk not in self.ws:
                    sht = self.bases[key[:2]].clone()
                    sht.configure(scatter=key[2], chunk=key[3])
                    self.ws[wk] = [sht, 0, 1]
                self.ws[wk][1] = max(self.ws[wk][1], max(self.sets[i][0].numel() for i in idx))
                self.ws[wk][2] = max(self.ws[wk][2], len(idx))
        for sht, n, k in self.ws.values():
            sht.reserve(n, k, weight_sets=False)
        self.call_out = [torch.empty((len(idx), key[0] * (key[0] + 1) // 2), dtype=torch.complex128,
                                     device=self.device) for key, idx in self.calls]
        self.call_status = [torch.zeros(len(idx), dtype=torch.int32, device=self.device)
                            for _, idx in self.calls]
        self.out = [None] * len(self.sets)
        for ci, (key, idx) in enumerate(self.calls):
            for r, i in enumerate(idx):
                self.out[i] = self.call_out[ci][r]
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.nstreams)]
        self.graph = None
        if graph:
            self._capture()

    def _call_cost(self, call):
        """Relative time of a call: binning and scatter ~ points x Nl; the table
        contraction streams 8 bytes per (l, m, node), ~10 point-orders each."""
        key, idx = call
        Nl, Nx = key[0], key[1]
        return sum(self.sets[i][0].numel() * Nl for i in idx) + 10 * (Nl * (Nl + 1) // 2) * Nx

    def _enqueue(self, origin):
        for j, calls in enumerate(self.assign):
            s = self.streams[j]
            s.wait_stream(origin)
            for ci in calls:
                key, idx = self.calls[ci]
                sht = self.ws[(j,) + key][0]
                sht.transform_many_device([self.sets[i][:3] for i in idx], out=self.call_out[ci],
                                          stream=s, status=self.call_status[ci])
        for s in self.streams:
            origin.wait_stream(s)

    def _capture(self):
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream(device=self.device)
        with torch.cuda.graph(self.graph, stream=cap):
            self._enqueue(cap)
        torch.cuda.synchronize(self.device)

    def run(self):
        """Replay the batch (one graph launch) on the current stream."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._enqueue(torch.cuda.current_stream(self.device))
        return self.out

    def run_eager(self):
        """The same work enqueued launch by launch (no graph)."""
        self._enqueue(torch.cuda.current_stream(self.device))
        return self.out

    def points_modes(self):
        return sum(s[0].numel() * (s[3] * (s[3] + 1) // 2) for s in self.sets)
