"""Many independent transforms of uneven size and band limit in one launch.

A batch is a list of point sets (device-resident theta, phi, w; each with its
own Nl, Nx).  BatchTransform prepares everything once -- resident tables per
(Nl, Nx), a tuned configuration per (Nl, size bucket), workspaces reserved to
size, output buffers -- and captures the whole batch as one HIP graph:
transforms are spread over several streams (longest first, onto the least
loaded stream), so small transforms overlap with large ones, and the graph
is replayed with a single launch.  Nothing in a replay allocates or
synchronizes.

Autotuner: for each (Nl, Nx, size bucket) present in a batch it times every
candidate (scatter engine, points per scatter work item) on a synthetic
uniform-sphere set of the bucket's size and keeps the fastest; results are
cached per process (and optionally in a JSON file).
"""
import json
import math
import os

import torch

from .sht import DirectSHT

CANDIDATES = (("matrix", 1024), ("matrix", 2048), ("matrix", 4096), ("vector", 2048))


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
            sht.configure(scatter=engine, chunk=chunk)
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
                 nx_factor=2, device=None, graph=True):
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
        # tuned configuration per set
        self.config = [self.tuner.choose(self.bases[(s[3], s[4])], s[0].numel()) for s in self.sets]
        # assignment to streams: longest first onto the least loaded stream
        cost = [self._cost(s) for s in self.sets]
        order = sorted(range(len(self.sets)), key=lambda i: -cost[i])
        self.nstreams = max(1, int(nstreams))
        load = [0.0] * self.nstreams
        self.assign = [[] for _ in range(self.nstreams)]
        for i in order:
            j = min(range(self.nstreams), key=lambda q: load[q])
            self.assign[j].append(i)
            load[j] += cost[i]
        # one workspace per (stream, table, config), reserved for its largest set
        self.ws = {}
        for j, idx in enumerate(self.assign):
            for i in idx:
                th, _, _, Nl, Nx = self.sets[i]
                key = (j, Nl, Nx) + tuple(self.config[i])
                if key not in self.ws:
                    sht = self.bases[(Nl, Nx)].clone()
                    sht.configure(scatter=self.config[i][0], chunk=self.config[i][1])
                    self.ws[key] = [sht, 0]
                self.ws[key][1] = max(self.ws[key][1], th.numel())
        for sht, n in self.ws.values():
            sht.reserve(n)
        self.out = [torch.empty(s[3] * (s[3] + 1) // 2, dtype=torch.complex128, device=self.device)
                    for s in self.sets]
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.nstreams)]
        self.graph = None
        if graph:
            self._capture()

    @staticmethod
    def _cost(s):
        th, _, _, Nl, Nx = s
        Nlm = Nl * (Nl + 1) // 2
        return th.numel() * Nl + 0.5 * Nlm * Nx  # scatter ~ N Nl, table stream ~ Nlm Nx

    def _enqueue(self, origin):
        for j, idx in enumerate(self.assign):
            s = self.streams[j]
            s.wait_stream(origin)
            for i in idx:
                th, ph, w, Nl, Nx = self.sets[i]
                sht = self.ws[(j, Nl, Nx) + tuple(self.config[i])][0]
                sht.transform_device(th, ph, w, out=self.out[i], stream=s)
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
