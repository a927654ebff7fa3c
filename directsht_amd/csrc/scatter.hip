// Stage (b), part 2: scatter of the sorted points into per-(segment, m)
// accumulators  v_k(m, seg) = sum_{p in seg} w_p h_k(t_p) e^{-i m phi_p},
// k = 0..3 -- the v's of sht/shared_interp_funcs.py:88-105 (and
// csht/sht_helper.c:199-240), without trigonometric calls per (point, m).
//
// Work decomposition: one wavefront per (chunk of `chunk` sorted points,
// m-tile of 64*K orders).  Lane t owns the K consecutive orders
// m = m0 + t*K .. m0 + t*K + K-1 and keeps their 8 real accumulators in
// registers while it walks the chunk point by point (every lane sees the same
// point, so the segment changes are wave-uniform and flushes never diverge).
//
// e^{-i m phi} per (point, lane) comes from a seed e1^(m0 + tK), e1 = e^{-i phi},
// followed by K-1 complex multiplications by e1.  The 64 seeds of a point are
// produced cooperatively: for a batch of 8 points, 8 lanes per point each
// compute 8 consecutive seeds (one binary power, then a chain of 7 products
// by e1^K) into LDS.  Everything is FP64: on gfx950 FP64 FMA runs at the
// non-packed FP32 rate and the FP64 recurrences need no reseeding.
//
// A run is a maximal stretch of one segment inside one chunk; each run's sums
// go to their own slot part[run][m][8] (no atomics), and the fold stage adds
// the runs of a segment in a fixed order, so the reduction is deterministic.
#include "common.h"
#include "plan.h"

namespace dsht {

struct cplx {
  double x, y;
};

__device__ inline cplx cmul(cplx a, cplx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

__device__ inline cplx cpow(cplx e, int n) {
  cplx r{1.0, 0.0};
  while (n > 0) {
    if (n & 1) r = cmul(r, e);
    e = cmul(e, e);
    n >>= 1;
  }
  return r;
}

constexpr int kBatch = 8;  // points per seed batch

template <int K>
__global__ __launch_bounds__(64) void k_scatter64(SortedPoints sp, const int* __restrict__ seg_start,
                                                  int nseg, int chunk, int Nl,
                                                  const int* __restrict__ run_base,
                                                  double* __restrict__ part,
                                                  int* __restrict__ run_seg) {
  const int c = blockIdx.x, mt = blockIdx.y, lane = threadIdx.x;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;
  const int m0 = mt * 64 * K;
  const int mlane = m0 + lane * K;

  __shared__ cplx seeds[kBatch][64];
  __shared__ double pdat[kBatch][6];
  __shared__ int pseg[kBatch];

  double acc[K][8];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[k][q] = 0.0;

  int run = run_base[c];
  int cur = sp.seg[b];

  auto flush = [&]() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int m = mlane + k;
      if (m < Nl) {
        double2* dst = reinterpret_cast<double2*>(part + ((int64_t)run * Nl + m) * 8);
        dst[0] = make_double2(acc[k][0], acc[k][1]);
        dst[1] = make_double2(acc[k][2], acc[k][3]);
        dst[2] = make_double2(acc[k][4], acc[k][5]);
        dst[3] = make_double2(acc[k][6], acc[k][7]);
      }
    }
    if (lane == 0 && mt == 0) run_seg[run] = cur;
  };

  for (int pb = b; pb < e; pb += kBatch) {
    const int nb = min(kBatch, e - pb);
    {
      const int j = lane >> 3, r = lane & 7;
      if (j < nb) {
        const cplx e1{sp.ere[pb + j], sp.eim[pb + j]};
        const cplx eK = cpow(e1, K);
        const cplx e8K = cpow(eK, 8);
        cplx z = cpow(e8K, 8 * mt + r);  // e1^(m0 + 8rK), m0 = 64K mt
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          seeds[j][8 * r + q] = z;
          z = cmul(z, eK);
        }
      }
      if (lane < nb) {
        pdat[lane][0] = sp.ere[pb + lane];
        pdat[lane][1] = sp.eim[pb + lane];
        pdat[lane][2] = sp.h0[pb + lane];
        pdat[lane][3] = sp.h1[pb + lane];
        pdat[lane][4] = sp.h2[pb + lane];
        pdat[lane][5] = sp.h3[pb + lane];
        pseg[lane] = sp.seg[pb + lane];
      }
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

int launch_scatter(Plan& pl, int64_t Np, hipStream_t s) {
  constexpr int K = 8;
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  const int nmt = (pl.Nl + 64 * K - 1) / (64 * K);
  dim3 grid(nchunks, nmt);
  k_scatter64<K><<<grid, 64, 0, s>>>(pl.ws.sp, pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                     pl.ws.run_base, pl.ws.part, pl.ws.run_seg);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
