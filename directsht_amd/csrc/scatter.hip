// Stage (b), part 2: scatter of the sorted points into per-(segment, m)
// accumulators  v_k(m, seg) = sum_{p in seg} w_p h_k(t_p) e^{-i m phi_p},
// k = 0..3 -- the v's of sht/shared_interp_funcs.py:88-105 (and
// csht/sht_helper.c:199-240), without trigonometric calls per (point, m).
//
// Work decomposition: one wavefront per (chunk of `chunk` sorted points,
// m-tile of 64*K orders).  Lane t owns the K consecutive orders
// m = m0 + t*K .. m0 + t*K + K-1 and keeps their 8 real accumulators in
// registers while it walks the chunk point by point (every lane sees the same
// point, so segment changes are wave-uniform and flushes never diverge).
//
// e^{-i m phi} per (point, lane) comes from a seed e1^(m0 + tK), e1 = e^{-i phi},
// followed by K-1 complex multiplications by e1.  The 64 seeds of a point are
// produced cooperatively in FP64: for a batch of 8 points, 8 lanes per point
// each form e1^(m0 + 8rK) by repeated squaring and then a chain of 7 products
// by e1^K, into LDS.
//
// Two arithmetic variants:
//   FP64  everything in FP64 (DSHT_FP64 plans); FP64 recurrences drift by
//         ~1e-16 per step and need no reseeding.
//   FP32  seeds rounded to FP32, K-1 FP32 rotation steps (phase error
//         <~ K * 6e-8), and two-level FP32 accumulation: a fresh inner sum per
//         batch of 8 points, added into an outer sum per run (DSHT_FP32 plans;
//         on gfx950 the non-packed FP32 VALU loop issues ~1.8x faster than
//         FP64).
//
// A run is a maximal stretch of one segment inside one chunk; each run's sums
// go to their own slot part[run][k][m][2] (v_k, complex), with no atomics,
// and the fold stage adds the runs of a segment in a fixed order, so the
// reduction is deterministic.
#include "common.h"
#include "plan.h"

namespace dsht {

struct cplx {
  double x, y;
};

__device__ inline cplx cmul(cplx a, cplx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline cplx csq(cplx a) { return {a.x * a.x - a.y * a.y, 2.0 * a.x * a.y}; }

__device__ inline cplx cpow(cplx e, int n) {
  cplx r{1.0, 0.0};
  bool first = true;
  while (n > 0) {
    if (n & 1) {
      r = first ? e : cmul(r, e);
      first = false;
    }
    n >>= 1;
    if (n) e = csq(e);
  }
  return r;
}

constexpr int kBatch = 8;  // points per seed batch

// FP64 seeds of one batch: lane (j = lane/8, r = lane%8) writes
// seeds[j][8r + q] = e1_j^(m0 + (8r + q) K), q = 0..7.
template <int K, typename S, typename Store>
__device__ inline void make_seeds(const PointRec* __restrict__ rec, int pb, int nb, int mt,
                                  Store store) {
  const int lane = threadIdx.x, j = lane >> 3, r = lane & 7;
  if (j < nb) {
    const cplx e1{rec[pb + j].ere, rec[pb + j].eim};
    cplx eK = e1;
#pragma unroll
    for (int k = 1; k < K; k <<= 1) eK = csq(eK);  // e1^K (K a power of two)
    cplx e8K = csq(csq(csq(eK)));                  // e1^(8K)
    cplx z = cpow(e8K, 8 * mt + r);                // e1^(64K mt + 8K r)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      store(j, 8 * r + q, z);
      z = cmul(z, eK);
    }
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_scatter64(const PointRec* __restrict__ rec,
                                                  const int* __restrict__ seg_start, int nseg,
                                                  int chunk, int Nl,
                                                  const int* __restrict__ run_base,
                                                  const double* __restrict__ wsort,
                                                  double* __restrict__ part,
                                                  int* __restrict__ run_seg) {
  const int c = blockIdx.x, mt = blockIdx.y, lane = threadIdx.x;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;
  const int mlane = mt * 64 * K + lane * K;

  __shared__ cplx seeds[kBatch][64];
  __shared__ double pdat[kBatch][6];
  __shared__ int pseg[kBatch];

  double acc[K][8];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[k][q] = 0.0;

  int run = run_base[c];
  int cur = rec[b].seg;

  auto flush = [&]() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int m = mlane + k;
      if (m < Nl) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          *reinterpret_cast<double2*>(part + (((int64_t)run * 4 + q) * Nl + m) * 2) =
              make_double2(acc[k][2 * q], acc[k][2 * q + 1]);
      }
    }
    if (lane == 0 && mt == 0) run_seg[run] = cur;
  };

  for (int pb = b; pb < e; pb += kBatch) {
    const int nb = min(kBatch, e - pb);
    make_seeds<K, cplx>(rec, pb, nb, mt, [&](int j, int t, cplx z) { seeds[j][t] = z; });
    if (lane < nb) {
      const PointRec& pr = rec[pb + lane];
      const double ws = wsort ? wsort[pb + lane] : 1.0;
      pdat[lane][0] = pr.ere;
      pdat[lane][1] = pr.eim;
      pdat[lane][2] = ws * pr.h0;
      pdat[lane][3] = ws * pr.h1;
      pdat[lane][4] = ws * pr.h2;
      pdat[lane][5] = ws * pr.h3;
      pseg[lane] = pr.seg;
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
      const int sj = pseg[j];
      if (sj != cur) {
        flush();
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int q = 0; q < 8; ++q) acc[k][q] = 0.0;
        ++run;
        cur = sj;
      }
      cplx z = seeds[j][lane];
      const double er = pdat[j][0], ei = pdat[j][1];
      const double h0 = pdat[j][2], h1 = pdat[j][3], h2 = pdat[j][4], h3 = pdat[j][5];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        acc[k][0] = fma(h0, z.x, acc[k][0]);
        acc[k][1] = fma(h0, z.y, acc[k][1]);
        acc[k][2] = fma(h1, z.x, acc[k][2]);
        acc[k][3] = fma(h1, z.y, acc[k][3]);
        acc[k][4] = fma(h2, z.x, acc[k][4]);
        acc[k][5] = fma(h2, z.y, acc[k][5]);
        acc[k][6] = fma(h3, z.x, acc[k][6]);
        acc[k][7] = fma(h3, z.y, acc[k][7]);
        const double zr = fma(z.x, er, -z.y * ei);
        const double zi = fma(z.x, ei, z.y * er);
        z.x = zr;
        z.y = zi;
      }
    }
    __syncthreads();
  }
  flush();
}

template <int K>
__global__ __launch_bounds__(64) void k_scatter32(const PointRec* __restrict__ rec,
                                                  const int* __restrict__ seg_start, int nseg,
                                                  int chunk, int Nl,
                                                  const int* __restrict__ run_base,
                                                  const double* __restrict__ wsort,
                                                  float* __restrict__ part,
                                                  int* __restrict__ run_seg) {
  const int c = blockIdx.x, mt = blockIdx.y, lane = threadIdx.x;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;
  const int mlane = mt * 64 * K + lane * K;

  __shared__ float2 seeds[kBatch][64];
  __shared__ float pdat[kBatch][6];
  __shared__ int pseg[kBatch];

  float acc[K][8], inn[K][8];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[k][q] = inn[k][q] = 0.f;

  int run = run_base[c];
  int cur = rec[b].seg;

  auto merge = [&]() {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        acc[k][q] += inn[k][q];
        inn[k][q] = 0.f;
      }
  };
  auto flush = [&]() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int m = mlane + k;
      if (m < Nl) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          *reinterpret_cast<float2*>(part + (((int64_t)run * 4 + q) * Nl + m) * 2) =
              make_float2(acc[k][2 * q], acc[k][2 * q + 1]);
      }
    }
    if (lane == 0 && mt == 0) run_seg[run] = cur;
  };

  for (int pb = b; pb < e; pb += kBatch) {
    const int nb = min(kBatch, e - pb);
    make_seeds<K, float2>(rec, pb, nb, mt,
                          [&](int j, int t, cplx z) { seeds[j][t] = make_float2((float)z.x, (float)z.y); });
    if (lane < nb) {
      const PointRec& pr = rec[pb + lane];
      const double ws = wsort ? wsort[pb + lane] : 1.0;
      pdat[lane][0] = (float)pr.ere;
      pdat[lane][1] = (float)pr.eim;
      pdat[lane][2] = (float)(ws * pr.h0);
      pdat[lane][3] = (float)(ws * pr.h1);
      pdat[lane][4] = (float)(ws * pr.h2);
      pdat[lane][5] = (float)(ws * pr.h3);
      pseg[lane] = pr.seg;
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
      const int sj = pseg[j];
      if (sj != cur) {
        merge();
        flush();
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int q = 0; q < 8; ++q) acc[k][q] = 0.f;
        ++run;
        cur = sj;
      }
      float2 z = seeds[j][lane];
      const float er = pdat[j][0], ei = pdat[j][1];
      const float h0 = pdat[j][2], h1 = pdat[j][3], h2 = pdat[j][4], h3 = pdat[j][5];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        inn[k][0] = fmaf(h0, z.x, inn[k][0]);
        inn[k][1] = fmaf(h0, z.y, inn[k][1]);
        inn[k][2] = fmaf(h1, z.x, inn[k][2]);
        inn[k][3] = fmaf(h1, z.y, inn[k][3]);
        inn[k][4] = fmaf(h2, z.x, inn[k][4]);
        inn[k][5] = fmaf(h2, z.y, inn[k][5]);
        inn[k][6] = fmaf(h3, z.x, inn[k][6]);
        inn[k][7] = fmaf(h3, z.y, inn[k][7]);
        const float zr = fmaf(z.x, er, -z.y * ei);
        const float zi = fmaf(z.x, ei, z.y * er);
        z.x = zr;
        z.y = zi;
      }
    }
    merge();
    __syncthreads();
  }
  flush();
}

int launch_scatter(Plan& pl, int64_t Np, const double* wsort, hipStream_t s) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  if (uses_spread(pl)) return launch_scatter_spread(pl, Np, wsort, wsort ? 1 : 0, s);
  if (pl.precision == DSHT_FP32 &&
      (pl.scatter_now == DSHT_SCATTER_MATRIX || pl.scatter_now == DSHT_SCATTER_SPREAD))
    return launch_scatter_mfma(pl, Np, wsort, false, s);
  if (pl.precision == DSHT_FP32) {
    constexpr int K = 8;
    dim3 grid(nchunks, (pl.Nl + 64 * K - 1) / (64 * K));
    k_scatter32<K><<<grid, 64, 0, s>>>(pl.ws.rec, pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                       pl.ws.run_base, wsort, (float*)pl.ws.part, pl.ws.run_seg);
  } else {
    constexpr int K = 8;
    dim3 grid(nchunks, (pl.Nl + 64 * K - 1) / (64 * K));
    k_scatter64<K><<<grid, 64, 0, s>>>(pl.ws.rec, pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                       pl.ws.run_base, wsort, (double*)pl.ws.part, pl.ws.run_seg);
  }
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
