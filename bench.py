#!/usr/bin/env python
"""Benchmark of the direct SHT hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the configuration the headline metric is quoted on): plain
DirectSHT(Nl=1000, Nx=2048, dflt_type='float32')(theta, phi, w) on 1e7 points
in the cap cos(theta) in [0.3, 1], density ~ exp(g - sigma^2/2) with g a
Gaussian random field (C_l ~ (l+10)^-2, sigma = 0.5), w = 1.  Inputs are
generated once on the GPU (not timed) and stay resident in HBM; a step is one
full transform (binning + sort, scatter, fold, contraction, polar rotation)
returning healpy-ordered complex128 alm in HBM.

Printed JSON line (stdout):  value = N * Nlm / seconds per step
(points*modes/s); roofline of the dominant kernel from HIP events recorded
on its stream during the timed steps; cpu_baseline = the reference's own C
(csht, compiled into oracle/_ref) on a bounded sample of the same workload.
Also (same line): "batched" -- many weight sets on shared points and an
uneven batch of point sets, batched vs plain single calls, with errors --
"serving" -- p50/p99 latency and aggregate throughput of the transform
server under Poisson arrivals -- and "downstream" -- power spectra, the
mode-coupling matrix and the Wigner-3j table on the device.  Progress goes
to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "points·modes/s and alm/s at Nl=1000,1e7 pts; batched vs plain; p50/p99 latency"
PEAK_FP32_TFLOPS = 157.3   # MI355X published dense FP32 (vector and matrix)
PEAK_FP16_TFLOPS = 2500.0  # MI355X published dense FP16 matrix
PEAK_FP64_TFLOPS = 78.6    # MI355X published dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
# HBM bytes per launch of each kernel at the default workload, measured by
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs scaled by a copy kernel of
# known size; the summary holding them is read when the workload matches.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_plain_c2.md")


def read_traffic():
    """{kernel: (fetch_bytes, write_bytes)} from the PMC summary's
    'traffic <kernel> <fetch> <write>' lines, or {} when absent."""
    out = {}
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic":
                    out[parts[1]] = (float(parts[2]), float(parts[3]))
    except OSError:
        pass
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        return ""


def cpu_baseline(theta, phi, w, Nl, Nx, N_full, sizes=(20000, 100000), reps=3, xmax=0.75):
    """The reference's csht do_transform (OpenMP C, compiled into oracle/_ref)
    on samples of the workload's points: equatorial points directly, polar
    points after the 90-degree rotation; tables built before timing.  Each
    sample size runs `reps` times (median); the time per transform is fitted
    as t = t_fixed + n * t_point (csht contracts every bin for each call,
    a fixed O(Nlm Nx) cost, and scatters O(n Nl)), and the full workload's
    rate is N_full * Nlm / (t_fixed + N_full * t_point).  The O(Nl^3) alm
    rotation (healpy in the reference) is not included."""
    import numpy as np

    from oracle import cref, refalg
    if not cref.available():
        return None
    log(f"cpu baseline: csht make_table Nl={Nl} Nx={Nx} (not timed)")
    c = cref.CRef(Nl, Nx, xmax)
    med = {}
    for n in sizes:
        t, p, ww = theta[:n], phi[:n], w[:n]
        x = np.abs(np.cos(t))
        pol, equ = x >= xmax, x < xmax
        tp, pp = refalg.rotate_points(t[pol], p[pol])
        runs = []
        for r in range(reps):
            log(f"cpu baseline: csht on {n} points, run {r + 1}/{reps}")
            t0 = time.perf_counter()
            c(t[equ], p[equ], ww[equ])
            c(tp, pp, ww[pol])
            runs.append(time.perf_counter() - t0)
        med[n] = float(np.median(runs))
    n0, n1 = sizes
    t_point = (med[n1] - med[n0]) / (n1 - n0)
    t_fixed = max(0.0, med[n0] - n0 * t_point)
    t_full = t_fixed + N_full * t_point
    Nlm = Nl * (Nl + 1) // 2
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": N_full * Nlm / t_full,
        "unit": "points*modes/s",
        "cores": cores,
        "nproc": os.cpu_count(),
        "cpu_model": _cpu_model(),
        "kind": "reference",
        "sample": (f"csht/sht_helper.c do_transform (its csht/sht.py driver, both branches, "
                   f"tables prebuilt, alm rotation excluded) on the first {n0} and {n1} points "
                   f"of the same workload (Nl={Nl}, Nx={Nx}), median of {reps} runs each: "
                   f"{med[n0]:.2f} s and {med[n1]:.2f} s; fitted fixed {t_fixed:.2f} s + "
                   f"{t_point * 1e6:.1f} us/point, so {t_full:.1f} s for all {N_full} points"),
        "median_s": {str(k): v for k, v in med.items()},
        "fixed_s": t_fixed,
        "per_point_us": t_point * 1e6,
        "full_workload_s": t_full,
    }


def cpu_numpy_c1(reps=3):
    """The numpy restatement of the reference's numpy path (oracle.refalg,
    FP64) on the full C1 workload (Nl=256, Nx=512, 1e5 uniform points), tables
    built before timing; median of reps runs on one process."""
    import numpy as np

    from oracle import refalg
    rng = np.random.default_rng(1)
    n = 100_000
    t = np.arccos(rng.uniform(-1, 1, n))
    p = rng.uniform(0, 2 * np.pi, n)
    w = rng.uniform(0.5, 1.5, n)
    T = refalg.make_tables(256, 512, 0.75)
    runs = []
    for r in range(reps):
        log(f"cpu baseline: numpy oracle on C1, run {r + 1}/{reps}")
        t0 = time.perf_counter()
        refalg.direct_sht(T, t, p, w)
        runs.append(time.perf_counter() - t0)
    med = float(np.median(runs))
    return {"value": n * 32896 / med, "unit": "points*modes/s", "median_s": med, "runs": len(runs),
            "cores": 1, "kind": "port",
            "sample": "oracle.refalg.direct_sht (numpy restatement of sht/sht.py:112-283, "
                      "FP64) on the whole C1 workload: Nl=256, Nx=512, 1e5 uniform points"}


def method_error(Nl, Nx, theta, phi, W, n, dflt_type="float32"):
    """Error of the device transform against the exact FP64 direct sum
    (oracle.exact, no spline) on the first n points: max |da| / max |a| and
    rms, worst over the weight rows W (S, N)."""
    import numpy as np
    import torch

    from directsht_amd.sht import DirectSHT
    from oracle import exact
    t = theta[:n].contiguous()
    p = phi[:n].contiguous()
    Wn = W[:, :n].contiguous()
    ex = exact.exact_alm(Nl, t.cpu().numpy(), p.cpu().numpy(), Wn.cpu().numpy())
    sht = DirectSHT(Nl, Nx, dflt_type=dflt_type)
    if Wn.shape[0] == 1:
        a = sht.transform_device(t, p, Wn[0].contiguous())[0][None, :]
    else:
        a = sht.transform_sets_device(t, p, Wn)[0]
    a = a.cpu().numpy()
    d = np.abs(a - ex)
    mx = float(np.max(d.max(axis=1) / np.abs(ex).max(axis=1)))
    rms = float(np.max(np.linalg.norm(a - ex, axis=1) / np.linalg.norm(ex, axis=1)))
    del sht
    torch.cuda.empty_cache()
    return {"max_rel": mx, "rms_rel": rms, "points": n, "sets": int(Wn.shape[0]),
            "vs": "exact FP64 direct sum (oracle.exact)", "plan": dflt_type}


def _timed(fn, reps, sync):
    """Mean seconds per call of fn() over reps calls (after one untimed call)."""
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    sync()
    return (time.perf_counter() - t0) / reps


def _err(a, ref):
    """max over sets of (max-rel, rms-rel) of a vs ref, both (S, Nlm)."""
    d = (a - ref).abs()
    mx = (d.amax(dim=1) / ref.abs().amax(dim=1)).max()
    rms = (d.norm(dim=1) / ref.norm(dim=1)).max()
    return {"max_rel": float(mx), "rms_rel": float(rms)}


def bench_weight_sets(reps=3, headline=False, S=64):
    """Many weight sets on shared points.  Default (the weight-set
    configuration): 2e6 uniform points, Nl=768, Nx=1536.  headline=True: the
    1e7 lognormal-cap points of the plain workload at Nl=1000, Nx=2048.  S
    weight sets: set 0 constant, the others lognormal-modulated with
    distinct seeds; FP32.  Plain = S single-set transforms; batched = one
    dsht_transform_sets call (binning once, one table stream, fp16x3 MFMA
    contraction; for the default configuration also with the FP32 GEMV
    contraction per set).  Errors against the FP64 plan per set, and against
    the exact direct sum on a subsample."""
    import torch

    from directsht_amd import mocks
    from directsht_amd.sht import DirectSHT
    if headline:
        Nl, Nx, N = 1000, 2048, 10_000_000
        th, ph, _ = mocks.lognormal_cap_points(N, seed=2)
        what = f"{N} lognormal cap points (the plain workload's)"
    else:
        Nl, Nx, N = 768, 1536, 2_000_000
        th, ph, _ = mocks.uniform_sphere(N, seed=3)
        what = f"{N} uniform points"
    log(f"weight sets: {what}, {S} weight sets, Nl={Nl}")
    W = mocks.lognormal_weight_sets(th, ph, S, seed0=300)
    sync = torch.cuda.synchronize
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    sht.reserve(N, S)
    out1 = torch.empty(sht.Nlm, dtype=torch.complex128, device="cuda")

    def plain():
        for s in range(S):
            sht.transform_device(th, ph, W[s], out=out1)

    outS = torch.empty((S, sht.Nlm), dtype=torch.complex128, device="cuda")
    t_plain = _timed(plain, reps, sync)
    t_mat = _timed(lambda: sht.transform_sets_device(th, ph, W, out=outS), reps, sync)
    a_mat = outS.clone()
    t_vec = a_vec = None
    if not headline:
        sht.configure(contract="vector")
        t_vec = _timed(lambda: sht.transform_sets_device(th, ph, W, out=outS), reps, sync)
        a_vec = outS.clone()
    del sht, outS
    torch.cuda.empty_cache()
    s64 = DirectSHT(Nl, Nx, dflt_type="float64")
    a64, _ = s64.transform_sets_device(th, ph, W)
    sync()
    Nlm = Nl * (Nl + 1) // 2
    pm = N * Nlm * S
    res = {"workload": f"{what} x {S} weight sets (1 constant, {S - 1} lognormal "
                       f"sigma=0.5, seeds 301..), Nl={Nl}, Nx={Nx}, FP32",
           "plain_ms": t_plain * 1e3, "batched_ms": t_mat * 1e3,
           "plain_points_modes_per_s": pm / t_plain, "batched_points_modes_per_s": pm / t_mat,
           "speedup": t_plain / t_mat,
           "error_vs_fp64": {"mfma_fp16x3": _err(a_mat, a64)}}
    if a_vec is not None:
        res["batched_vector_contract_ms"] = t_vec * 1e3
        res["error_vs_fp64"]["fp32_gemv"] = _err(a_vec, a64)
    del s64, a64, a_mat, a_vec
    torch.cuda.empty_cache()
    if not headline:
        log("weight sets: error against the exact direct sum (1000 points, 2 sets)")
        res["method_error_vs_exact"] = method_error(Nl, Nx, th, ph, W[:2], 1000)
    del W, th, ph
    torch.cuda.empty_cache()
    return res


def bench_uneven(nsets=200, reps=3):
    """Uneven batch: 200 independent uniform point sets, sizes log-uniform in
    [1e4, 5e6], Nl from {128, 256, 512, 1024}, Nx = 2 Nl, w = U[0.5, 1.5].
    Plain = one DirectSHT call per set (each returns when its alm is ready);
    batched = the autotuned BatchTransform graph (one launch)."""
    import numpy as np
    import torch

    from directsht_amd.batch import BatchTransform
    from directsht_amd.sht import DirectSHT
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(5e6), nsets)).astype(np.int64)
    nls = rng.choice([128, 256, 512, 1024], nsets)
    log(f"uneven batch: {nsets} sets, {int(sizes.sum())} points")
    g = torch.Generator(device="cuda").manual_seed(4)
    sets = []
    for n, nl in zip(sizes, nls):
        n = int(n)
        th = torch.acos(2 * torch.rand(n, generator=g, device="cuda", dtype=torch.float64) - 1)
        ph = 2 * math.pi * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
        w = 0.5 + torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
        sets.append((th, ph, w, int(nl)))
    sync = torch.cuda.synchronize
    t0 = time.perf_counter()
    bt = BatchTransform(sets, dflt_type="float32", nstreams=4)
    t_setup = time.perf_counter() - t0
    plans = {nl: bt.bases[(nl, 2 * nl)] for nl in set(int(x) for x in nls)}

    def plain():
        for th, ph, w, nl in sets:
            plans[nl].transform_device(th, ph, w)[1].item()  # the call returns with its result

    t_plain = _timed(plain, 1, sync)
    t_batch = _timed(bt.run, reps, sync)
    outs = [o.clone() for o in bt.run()]
    sync()
    worst = 0.0
    for i in range(0, nsets, 25):  # spot check against single calls
        th, ph, w, nl = sets[i]
        ref = plans[nl].transform_device(th, ph, w)[0]
        sync()
        worst = max(worst, float((outs[i] - ref).abs().max() / ref.abs().max()))
    pm = sum(int(n) * (int(nl) * (int(nl) + 1) // 2) for n, nl in zip(sizes, nls))
    res = {"workload": f"{nsets} uniform point sets, N log-uniform [1e4, 5e6] "
                       f"(total {int(sizes.sum())}), Nl in {{128,256,512,1024}}, Nx=2Nl, FP32",
           "plain_ms": t_plain * 1e3, "batched_ms": t_batch * 1e3, "speedup": t_plain / t_batch,
           "plain_points_modes_per_s": pm / t_plain, "batched_points_modes_per_s": pm / t_batch,
           "setup_s": t_setup, "calls": len(bt.calls), "streams": bt.nstreams,
           "max_rel_vs_single": worst}
    del bt, plans
    torch.cuda.empty_cache()
    big = next(i for i in range(nsets) if int(nls[i]) == 1024 and int(sizes[i]) >= 1000)
    th, ph, w, nl = sets[big]
    log("uneven: error against the exact direct sum (Nl=1024 set, 500 points)")
    res["method_error_vs_exact"] = method_error(1024, 2048, th, ph, w[None, :], 500)
    del sets
    torch.cuda.empty_cache()
    return res


def _event_ms(fn, reps):
    """Mean milliseconds of fn() on the current stream, from HIP events
    around reps calls (after one untimed call)."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_downstream(reps=20):
    """The consumers of alm on the device: power spectra of 64 alm sets at
    Nl=1000 (dsht_alm2cl, auto and all 2080 cross pairs), the mode-coupling
    matrix of MaskDeconvolution(564) (the reference notebook's Nl=500 plus 64
    buffer multipoles) and at Nl=2048 (dsht_mode_coupling), and the Wigner-3j
    table that MaskDeconvolution(564) builds, N=1127 (dsht_threej000_table);
    the lognormal mock generator at the reference's defaults (nside=2048,
    lmax=1000, 1e7 points) and its ring synthesis kernel.
    CPU: the reference's C 3j table (csht/threej000_helper.c, OpenMP) on the
    host's cores at N=1127."""
    import numpy as np
    import torch

    from directsht_amd import _lib
    from directsht_amd.mask_deconvolution import mode_coupling_device
    from directsht_amd.threej000 import threej000_table_device
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    out = {}
    Nl, S = 1000, 64
    Nlm = Nl * (Nl + 1) // 2
    g = torch.Generator(device="cuda").manual_seed(3)
    alm = torch.randn((S, Nlm), dtype=torch.complex128, device="cuda", generator=g)
    cl = torch.empty((S * (S + 1) // 2, Nl), dtype=torch.float64, device="cuda")
    for name, cross in (("alm2cl_64sets_Nl1000", 0), ("cross_spectra_64sets_Nl1000", 1)):
        ms = _event_ms(lambda: _lib.check(lib.dsht_alm2cl(Nl, S, alm.data_ptr(), Nlm, cl.data_ptr(),
                                                          cross, st), "dsht_alm2cl"), reps)
        nbytes = S * Nlm * 16
        out[name] = {"ms": ms, "alm_bytes": nbytes, "hbm_GBs": nbytes / ms / 1e6,
                     "hbm_frac": nbytes / ms / 1e6 / PEAK_HBM_GBS}
        if cross:
            # the Gram products on the FP64 matrix cores: 2 flop per (entry of
            # a 16 x 16 tile pair i <= j, l, term), 2 (l + 1) terms (m, re|im)
            # per l: 2 npairs 256 Nl (Nl + 1)
            nt = (S + 15) // 16
            mf = 2.0 * (nt * (nt + 1) // 2) * 256 * Nl * (Nl + 1)
            out[name].update({"fp64_mfma_tflops": mf / ms / 1e9,
                              "fp64_mfma_frac": mf / ms / 1e9 / PEAK_FP64_TFLOPS,
                              "bound": "max(HBM read, FP64 matrix products): both about 0.1 ms "
                                       "at peak for 64 sets at Nl=1000"})
        log(f"downstream {name}: {out[name]}")
    # CPU: the reference's alm2cl (sht/sht.py:95-109, restated in numpy by the
    # oracle) on 8 of the 64 sets, scaled to 64
    try:
        from oracle import refalg
        host = alm[:8].cpu().numpy()
        t0 = time.perf_counter()
        for a in host:
            refalg.alm2cl(a, Nl)
        dt = (time.perf_counter() - t0) * 8
        out["alm2cl_64sets_Nl1000"]["cpu_baseline"] = {
            "ms": dt * 1e3, "cores": 1, "kind": "port",
            "sample": "oracle.refalg.alm2cl (numpy) on 8 of the 64 sets, x8"}
    except Exception as e:  # reported, never fatal
        log(f"alm2cl cpu baseline failed: {e!r}")
    del alm, cl
    rng = np.random.default_rng(1)
    for nl in (564, 2048):
        W = torch.from_numpy(np.exp(-np.arange(2 * nl - 1) / 50.0) * rng.uniform(0.5, 1.5, 2 * nl - 1))
        W = W.cuda()
        M = torch.empty((nl, nl), dtype=torch.float64, device="cuda")
        ms = _event_ms(lambda: _lib.check(lib.dsht_mode_coupling(nl, W.data_ptr(), W.numel(),
                                                                 M.data_ptr(), st),
                                          "dsht_mode_coupling"), reps)
        terms = sum((l2 + 1) * (nl - l2) for l2 in range(nl))  # (l1 >= l2, l3) terms
        flop = 5.0 * terms
        out[f"mode_coupling_Nl{nl}"] = {"ms": ms, "terms": terms,
                                        "fp64_TFLOPs": flop / ms / 1e9,
                                        "fp64_frac": flop / ms / 1e9 / PEAK_FP64_TFLOPS}
        log(f"downstream mode_coupling Nl={nl}: {out[f'mode_coupling_Nl{nl}']}")
        del M
    N = 1127
    tab = threej000_table_device(N)
    ms = _event_ms(lambda: _lib.check(lib.dsht_threej000_table(N, tab.data_ptr(), st),
                                      "dsht_threej000_table"), 5)
    nbytes = tab.numel() * 8
    out["threej000_table_N1127"] = {"ms": ms, "bytes": nbytes, "hbm_GBs": nbytes / ms / 1e6,
                                    "hbm_frac": nbytes / ms / 1e6 / PEAK_HBM_GBS}
    del tab
    torch.cuda.empty_cache()
    try:
        from oracle import cref
        if cref.threej_available():
            t0 = time.perf_counter()
            tab = cref.threej000_table(N)
            dt = time.perf_counter() - t0
            out["threej000_table_N1127"]["cpu_baseline"] = {
                "ms": dt * 1e3, "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                "kind": "reference",
                "sample": "csht/threej000_helper.c make_table(1127), whole table"}
        # get_M at Nl=564 on the host from that table: the reference's sum
        # (sht/mask_deconvolution.py:183-191), vectorised over (l2, l3) per
        # row l1 with numpy, one thread
        if cref.threej_available():
            nl = (N + 1) // 2  # Wigner3j(2 nl - 1) is the table just built
            W = np.exp(-np.arange(2 * nl - 1) / 50.0)
            l2 = np.arange(nl)[:, None]
            l3 = np.arange(2 * nl - 1)[None, :]
            t0 = time.perf_counter()
            M = np.zeros((nl, nl))
            for l1 in range(nl):
                a_ = np.stack(np.broadcast_arrays(np.full_like(l2, l1), l2, l3))
                srt = np.sort(a_, axis=0)
                k3, k2, k1 = srt[0], srt[1], srt[2]
                tj = tab[(k1 * (k1 + 1) * (k1 + 2)) // 6 + (k2 * (k2 + 1)) // 2 + k3]
                M[l1] = ((2 * l3 + 1) * W[None, :] * tj * tj).sum(axis=1) * (2 * np.arange(nl) + 1)
            M /= 4 * np.pi
            dtm = time.perf_counter() - t0
            out["mode_coupling_Nl564"]["cpu_baseline"] = {
                "ms": (dtm + dt) * 1e3, "kind": "reference+port",
                "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                "sample": f"csht/threej000_helper.c make_table(1127) on those threads "
                          f"{dt:.2f} s + numpy get_M from it on one thread {dtm:.2f} s, "
                          f"whole matrix"}
            del tab
    except Exception as e:  # reported, never fatal
        log(f"3j / mode-coupling cpu baseline failed: {e!r}")
    # lognormal mocks at the reference's defaults (LogNormalMocks(Npnt, nside=2048,
    # lmax=1000)): the ring synthesis kernel alone, then a whole catalogue
    from directsht_amd import mocks
    Nl = 1000
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    alm = mocks.gaussian_alm(0.01 * (10.0 / (np.arange(Nl) + 10.0)) ** 2, gen)
    x = torch.cos((torch.arange(6144, dtype=torch.float64, device="cuda") + 0.5) * (math.pi / 6144))
    F = torch.empty((6144, Nl), dtype=torch.complex128, device="cuda")
    ms = _event_ms(lambda: _lib.check(lib.dsht_synthesis_rings(Nl, alm.data_ptr(), 6144, x.data_ptr(),
                                                               F.data_ptr(), st),
                                      "dsht_synthesis_rings"), 5)
    iters = 6144 * Nl * (Nl + 1) // 2
    out["synthesis_rings_Nl1000_6144rings"] = {"ms": ms, "ring_modes": iters,
                                               "fp64_TFLOPs": 6.0 * iters / ms / 1e9}
    del F
    lm = mocks.LogNormalMocks(10_000_000, nside=2048, lmax=1000)
    lm(seed=1, verbose=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, _, _ = lm(seed=2, verbose=False)
    out["lognormal_mocks_1e7_nside2048"] = {"s": time.perf_counter() - t0, "points": int(t.size),
                                            "grid": "6144 x 8192 pixels"}
    log(f"downstream mocks: {out['lognormal_mocks_1e7_nside2048']}")
    out["reference_notebook_s"] = {
        "MaskDeconvolution_564": 150.0,
        "source": "notebooks/analyzing_mocks.ipynb cell 17 output (Wigner3j(1127) + get_M, "
                  "numba + Python loops, the authors' machine)"}
    log(f"downstream 3j table: {out['threej000_table_N1127']}")
    return out


def bench_serving(nreq=2100, warm=100, rate=50.0):
    """Serving loop: Poisson arrivals (rate/s) of requests with N log-uniform
    in [1e4, 5e7] and Nl in {256, 512, 1024, 1536} (Nx = 2 Nl), FP32 resident
    tables; inputs are slices of a device-resident pool of uniform points;
    latency = submit -> alm on the host."""
    import numpy as np
    import torch

    from directsht_amd import mocks
    from directsht_amd.serve import TransformServer
    rng = np.random.default_rng(5)
    Nmax = 50_000_000
    th, ph, w = mocks.uniform_sphere(Nmax, seed=5)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(Nmax), nreq)).astype(np.int64)
    nls = rng.choice([256, 512, 1024, 1536], nreq)
    offs = [int(rng.integers(0, Nmax - n + 1)) for n in sizes]
    gaps = rng.exponential(1.0 / rate, nreq)
    log(f"serving: {nreq} requests ({warm} warm-up), Poisson {rate}/s")
    with TransformServer(max_inflight=4, dflt_type="float32") as srv:
        futs = []
        t_next = time.perf_counter()
        for i in range(nreq):
            t_next += gaps[i]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            a, n = offs[i], int(sizes[i])
            futs.append(srv.submit(th[a:a + n], ph[a:a + n], w[a:a + n], Nl=int(nls[i])))
            if (i + 1) % 500 == 0:
                log(f"serving: {i + 1} submitted")
        res = {}
        for i, f in enumerate(futs):
            a = f.result(timeout=600)
            if i >= warm and (i - warm) % 100 == 0:
                res[i] = a
        del futs
        done = srv.completed[warm:]
        srv.completed = done
        st = srv.stats()
        # capacity: the next requests of the same stream submitted all at
        # once (no arrival gaps), aggregate rate from first submit to last
        # result
        nsat = min(200, nreq)
        log(f"serving: {nsat} requests submitted at once (saturated)")
        srv.completed = []
        t0 = time.perf_counter()
        sfut = [srv.submit(th[offs[i]:offs[i] + int(sizes[i])], ph[offs[i]:offs[i] + int(sizes[i])],
                           w[offs[i]:offs[i] + int(sizes[i])], Nl=int(nls[i])) for i in range(nsat)]
        for f in sfut:
            f.result(timeout=600)
        t_sat = time.perf_counter() - t0
        pm_sat = sum(int(sizes[i]) * (int(nls[i]) * (int(nls[i]) + 1) // 2) for i in range(nsat))
        st["saturated"] = {"requests": nsat, "wall_s": t_sat,
                           "points_modes_per_s": pm_sat / t_sat,
                           "p50_ms": srv.stats()["p50_ms"], "p99_ms": srv.stats()["p99_ms"]}
    st["workload"] = (f"Poisson {rate}/s, {nreq - warm} timed requests after {warm} warm-up, "
                      f"N log-uniform [1e4, 5e7], Nl in {{256,512,1024,1536}}, Nx=2Nl, FP32, "
                      f"device-resident inputs, alm returned to the host")
    # served results against FP64 plans on the same points (every 100th
    # request), and the method error against the exact direct sum
    from directsht_amd.sht import DirectSHT
    worst = {"max_rel": 0.0, "rms_rel": 0.0}
    plans = {}
    for i in range(warm, nreq, 100):
        nl, a, n = int(nls[i]), offs[i], int(sizes[i])
        if nl not in plans:
            plans[nl] = DirectSHT(nl, 2 * nl, dflt_type="float64")
        ref = plans[nl].transform_device(th[a:a + n], ph[a:a + n], w[a:a + n])[0].cpu().numpy()
        got = np.asarray(res[i])
        worst["max_rel"] = max(worst["max_rel"],
                               float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))))
        worst["rms_rel"] = max(worst["rms_rel"],
                               float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
    worst["requests"] = len(range(warm, nreq, 100))
    st["served_error_vs_fp64"] = worst
    del plans
    torch.cuda.empty_cache()
    log("serving: error against the exact direct sum (Nl=1536, 300 points)")
    st["method_error_vs_exact"] = method_error(1536, 3072, th, ph, w[None, :], 300)
    del th, ph, w
    torch.cuda.empty_cache()
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--npoints", type=float, default=1e7)
    ap.add_argument("--cpu-sample", type=int, default=100000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--skip-batched", action="store_true", help="skip the batched workloads")
    ap.add_argument("--scatter", choices=("auto", "matrix", "vector", "spread"), default=None,
                    help="scatter engine of the plain workload (default: the plan's)")
    ap.add_argument("--chunk", type=int, default=None,
                    help="sorted points per scatter work item of the plain workload")
    ap.add_argument("--skip-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--serve-requests", type=int, default=2100)
    ap.add_argument("--skip-downstream", action="store_true",
                    help="skip the spectra / mode-coupling / 3j timings")
    ap.add_argument("--only", choices=("weight_sets", "weight_sets_headline",
                                       "weight_sets_headline_16", "uneven", "serving",
                                       "downstream"),
                    help="run only this workload and print its JSON (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after timing, run a 1 GiB device copy 3 times (the known-size "
                         "kernel that scales the PMC FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by points with an alm all-reduce; see the sharding design)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from directsht_amd import mocks
    from directsht_amd.sht import DirectSHT

    torch.cuda.set_device(0)
    if args.only:
        fn = {"weight_sets": bench_weight_sets, "uneven": bench_uneven,
              "weight_sets_headline": lambda: bench_weight_sets(headline=True, S=64),
              "weight_sets_headline_16": lambda: bench_weight_sets(headline=True, S=16),
              "downstream": bench_downstream,
              "serving": lambda: bench_serving(nreq=args.serve_requests,
                                               warm=min(100, args.serve_requests // 10))}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        if args.copy_calibration:
            src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
            dst = torch.empty_like(src)
            for _ in range(3):
                dst.copy_(src)
            torch.cuda.synchronize()
            log("copy calibration: 3 x 1 GiB device copies done")
        return
    Nl, Nx, N = 1000, 2048, int(args.npoints)
    Nlm = Nl * (Nl + 1) // 2
    log(f"generating {N} lognormal cap points")
    th, ph, w = mocks.lognormal_cap_points(N, seed=2)
    torch.cuda.synchronize()
    log("building tables")
    sht = DirectSHT(Nl, Nx, dflt_type="float32")
    if args.scatter or args.chunk:
        sht.configure(scatter=args.scatter, chunk=args.chunk)
    out = torch.empty(Nlm, dtype=torch.complex128, device="cuda")
    for _ in range(args.warmup):
        sht.transform_device(th, ph, w, out=out)
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    sht.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sht.transform_device(th, ph, w, out=out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stage_ms, ntimed = sht.get_timing()
    sht.set_timing(False)
    ms = (t1 - t0) * 1e3 / args.steps
    stage_avg = {k: v / max(ntimed, 1) for k, v in stage_ms.items()}
    log(f"{ms:.3f} ms/step; stages {stage_avg}")

    # Dominant kernel: the contraction (k_contract_otf + k_add_der).  FP32
    # plans hold no Y table: the kernel regenerates Y_lm at the nodes by the
    # FP64 Legendre recurrence (4 flop per (l, m, node)) and multiplies it
    # with the node coefficients on the FP32 matrix cores.  Algorithmic work
    # = the contraction's flop count, 8 per (l, m, node, branch) -- the
    # complex coefficients of Y and Y' -- over the 2 branches (equatorial,
    # polar): 16 Nlm Nx per transform, priced against the FP32 matrix peak.
    # Issued on the matrix cores: 16 columns (12 used: c, f, g x 4) x 2 flop
    # per (l, m, node) over the 2 * 2 ceil(Nx/4) padded nodes.
    npair = 2 * ((Nx + 3) // 4)
    con_s = stage_avg["contract"] * 1e-3
    con_tf = 16.0 * Nlm * Nx / con_s / 1e12
    con_mfma_tf = 32.0 * Nlm * 2 * npair / con_s / 1e12
    # The scatter (second largest stage).  The reference's algorithm does
    # 22 flop per (point, order m) -- four complex multiply-adds
    # w h_k(t) e^{-i m phi} into the accumulators (16) and the complex step
    # of e^{-i m phi} (6) -- 22 N Nl per transform.  The matrix engine
    # executes that form on the fp16 matrix cores (the real GEMM
    # [(k, m1, re|im) x point] . [point x (re|im, m2)], 32 flop per (point,
    # m), three times for the fp16 hi/lo split = 96 MFMA flop).  The
    # spreading engine computes the same sums as non-uniform FFTs (32
    # fixed-point grid updates per point; on grids up to 4096 cells with at
    # least 512 node pairs the node variant k_spread_nodes, one FFT of four
    # grids per (node, branch), the fold included; otherwise one FFT per run
    # and the fold after it), so 22 N Nl / t is only its
    # direct-sum-equivalent rate; its own bound is LDS (integer atomics, FFT
    # exchanges) and its HBM bytes are the 32-byte sorted records in and the
    # node coefficients (or run partials) out.
    scat_s = stage_avg["scatter"] * 1e-3
    engine = sht.scatter_engine(N)
    default_workload = (N == 10_000_000)
    traffic = read_traffic() if default_workload else {}

    def tr(name):
        t = traffic.get(name)
        return None if t is None else t[0] + t[1]

    alg_tf = 22.0 * N * Nl / scat_s / 1e12
    if engine == "matrix":
        mfma_tf = 96.0 * N * Nl / scat_s / 1e12
        roof_scatter = {
            "kernel": "k_scatter_mfma", "engine": "matrix (MFMA fp16 x3 split)", "bound": "mfma",
            "achieved": alg_tf, "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
            "frac": alg_tf / PEAK_FP16_TFLOPS, "traffic": tr("k_scatter_mfma"),
            "algorithmic": "22 flop per (point, m); 22 N Nl per transform",
            "mfma_instruction_tflops": mfma_tf, "mfma_instruction_frac": mfma_tf / PEAK_FP16_TFLOPS,
            "mfma_instruction_flop": "96 per (point, m): 32-flop real GEMM form, x3 for the "
                                     "fp16 hi/lo split"}
    else:
        node = 2 * ((Nx + 3) // 4) >= 512 and Nl <= 1024
        kname = "k_spread_nodes" if node else "k_scatter_spread"
        rec_bytes = 32.0 * N
        roof_scatter = {
            "kernel": kname, "engine": engine + (" (node variant, fold included)" if node else ""),
            "bound": "lds",
            "direct_sum_equivalent_tflops": alg_tf,
            "direct_sum_equivalent": "22 N Nl / t: the reference algorithm's flop count, "
                                     "not flop executed (non-uniform FFT per run)",
            "records": "32-byte sorted records in (N x 32 B)",
            "hbm_GBs_records_in": rec_bytes / scat_s / 1e9,
            "hbm_frac_records_in": rec_bytes / scat_s / 1e9 / PEAK_HBM_GBS,
            "traffic": tr(kname)}

    result = {
        "metric": METRIC,
        "value": N * Nlm / (ms * 1e-3),
        "unit": "points*modes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (lognormal cap points, seeded)",
        "config": {"workload": "plain DirectSHT Nl=1000 Nx=2048 xmax=0.75, 1e7 points, "
                               "cos(theta) in [0.3,1], lognormal density sigma=0.5, w=1, "
                               "FP32 contraction (FP64 binning and rotation)",
                   "Nl": Nl, "Nx": Nx, "npoints": N, "parallelism": "single GPU"},
        "alm_per_s": Nlm / (ms * 1e-3),
        "stage_ms": stage_avg,
        "roofline": {"kernel": "k_contract_otf + k_add_der", "bound": "mfma",
                     "engine": "FP32 matrix cores (v_mfma_f32_16x16x4f32), Y by FP64 recurrence",
                     "achieved": con_tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": con_tf / PEAK_FP32_TFLOPS, "traffic": tr("k_contract_otf"),
                     "algorithmic": "8 flop per (l, m, node, branch), 2 branches: 16 Nlm Nx per "
                                    "transform; no table is read",
                     "mfma_instruction_tflops": con_mfma_tf,
                     "mfma_instruction_frac": con_mfma_tf / PEAK_FP32_TFLOPS},
        "roofline_scatter": roof_scatter,
    }

    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 1 GiB device copies done")
        del src, dst

    # The drop-in call on host arrays, DirectSHT(...)(theta, phi, w) with
    # numpy in and out: the inputs' 240 MB go to the device and the alm's
    # 8 MB come back inside the call (not the bench's value, which starts
    # from device-resident inputs).
    log("host-inclusive: the numpy call")
    th_h, ph_h, w_h = (a.cpu().numpy() for a in (th, ph, w))
    sht(th_h, ph_h, w_h)
    torch.cuda.synchronize()
    reps_h = 3
    t0 = time.perf_counter()
    for _ in range(reps_h):
        sht(th_h, ph_h, w_h)
    host_ms = (time.perf_counter() - t0) * 1e3 / reps_h
    result["host_inclusive"] = {
        "ms": host_ms, "points_modes_per_s": N * Nlm / (host_ms * 1e-3), "reps": reps_h,
        "what": "DirectSHT(...)(numpy theta, phi, w) -> numpy alm: 3 x 80 MB of FP64 inputs "
                "to the device and 8 MB of alm back, pageable host memory, inside the call"}
    log(f"host-inclusive: {host_ms:.2f} ms per call")
    del th_h, ph_h, w_h

    if not args.no_accuracy:
        log("accuracy: FP64 plan on the same points")
        a32 = out.clone()
        del sht
        torch.cuda.empty_cache()
        s64 = DirectSHT(Nl, Nx, dflt_type="float64")
        a64, _ = s64.transform_device(th, ph, w)
        torch.cuda.synchronize()
        d = (a32 - a64).abs()
        result["accuracy_fp32_vs_fp64"] = {
            "max_rel": float(d.max() / a64.abs().max()),
            "rms_rel": float(d.norm() / a64.norm())}
        del s64
        torch.cuda.empty_cache()
        log("accuracy: error against the exact direct sum (1000 points)")
        result["method_error_vs_exact"] = method_error(Nl, Nx, th, ph, w[None, :], 1000)

    if not args.no_cpu_baseline:
        n = min(args.cpu_sample, N)
        sample = [a[:n].cpu().numpy() for a in (th, ph, w)]
        try:
            result["cpu_baseline"] = cpu_baseline(*sample, Nl, Nx, N, sizes=(n // 5, n))
        except Exception as e:  # reported, never fatal for the GPU number
            log(f"cpu baseline failed: {e!r}")
            result["cpu_baseline"] = None
        if result.get("cpu_baseline"):
            result["cpu_baseline"]["gpu_over_cpu"] = result["value"] / result["cpu_baseline"]["value"]
        try:
            c1 = cpu_numpy_c1()
            from directsht_amd.sht import DirectSHT as _D
            import numpy as _np
            rng = _np.random.default_rng(1)
            t1 = torch.from_numpy(_np.arccos(rng.uniform(-1, 1, 100_000))).cuda()
            p1 = torch.from_numpy(rng.uniform(0, 2 * _np.pi, 100_000)).cuda()
            w1 = torch.from_numpy(rng.uniform(0.5, 1.5, 100_000)).cuda()
            s1 = _D(256, 512, dflt_type="float64")
            g1 = _timed(lambda: s1.transform_device(t1, p1, w1), 20, torch.cuda.synchronize)
            c1["gpu_fp64_ms"] = g1 * 1e3
            c1["gpu_over_cpu"] = c1["median_s"] / g1
            result["cpu_baseline_c1_numpy"] = c1
            del s1
        except Exception as e:
            log(f"numpy C1 baseline failed: {e!r}")
    if not args.skip_batched:
        result["batched"] = {}
        for name, fn in (("weight_sets", bench_weight_sets), ("uneven", bench_uneven),
                         ("weight_sets_headline_16",
                          lambda: bench_weight_sets(headline=True, S=16)),
                         ("weight_sets_headline_64",
                          lambda: bench_weight_sets(headline=True, S=64))):
            try:
                result["batched"][name] = fn()
                log(f"{name}: {result['batched'][name]}")
            except Exception as e:  # reported, never fatal for the headline number
                log(f"{name} failed: {e!r}")
                result["batched"][name] = {"error": repr(e)}
    if not args.skip_serving:
        try:
            result["serving"] = bench_serving(nreq=args.serve_requests,
                                              warm=min(100, args.serve_requests // 10))
            log(f"serving: {result['serving']}")
        except Exception as e:
            log(f"serving failed: {e!r}")
            result["serving"] = {"error": repr(e)}
    if not args.skip_downstream:
        try:
            result["downstream"] = bench_downstream()
        except Exception as e:
            log(f"downstream failed: {e!r}")
            result["downstream"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
