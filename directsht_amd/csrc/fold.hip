// Stage (b), part 3: reduce the per-run partials of each segment and fold the
// four Hermite components onto the spline nodes and the two hemispheres.
//
// The reference contracts per bin j (sht/interp_funcs_py.py:46-60):
//   a = sum_j Y(x_j) v0(j) + dx Y'(x_j) v1(j) + Y(x_{j+1}) v2(j) + dx Y'(x_{j+1}) v3(j)
// and adds the x < 0 hemisphere with (-1)^(l-m) (sht/sht.py:185,279).  The
// same sum, regrouped by node n, is
//   a = sum_n Y(x_n) c(n) + Y'(x_n) d(n),
//   c(n) = v0(n) + v2(n-1),   d(n) = dx (v1(n) + v3(n-1)),
// and with s = c+ + c-, delta = c+ - c- the rows with even l-m contract s and
// the rows with odd l-m contract delta -- one table read for both hemispheres.
//
// The runs of a segment are contiguous in run order (runs are numbered by
// position in the sorted array); k_seg_runs records each segment's run range
// and k_fold sums them in order (deterministic).  k_fold: one thread per
// (order m, node pair np) for both branches; the 64 lanes of a wavefront take
// 64 consecutive m, so every partial it reads is one 512-byte (FP32) block,
// and each thread writes the whole 16-value row rhs[m][P][np][set][16] (64
// bytes, FP32) that the contraction reads (set = weight set or point set of a
// batched transform, one slot for a single transform):
//   branch b, offset 8b:  re: c(2np), c(2np+1), d(2np), d(2np+1);  im: same.
// It also records max(|c|, |d/dx|) of the set (as float bits, atomicMax) for
// the scaling of the matrix-core contraction.
//
// For the plain contraction, which streams Y only, the Y' term is rewritten
// with the derivative relation of the tables (sht/legendre_py.py:54-56,
// 111-117 in normalized form)
//   Y'_l(x) = [G_l Y_{l-1}(x) - l x Y_l(x)] / (1 - x^2),
//   G_l = sqrt((l^2 - m^2)(2l+1)/(2l-1)),
// so that  sum_n Y'_l(x_n) d_P(n) = G_l sum_n Y_{l-1}(x_n) g(n) - l sum_n Y_l(x_n) f_P(n)
// with f_P(n) = x_n d_P(n)/(1 - x_n^2) and g(n) = d_P(n)/(1 - x_n^2).  Row l-1
// has the other parity, so the rows of parity P are contracted with c_P, f_P
// and g = d_{1-P}/(1 - x^2) (for the row above them).  That fold writes the
// plain records: per (node pair, branch, m, P) [c | f | g][re, im][node 2np,
// 2np+1] at plain_rec (a pair's branch is one block, which the spreading
// scatter's node variant writes from one workgroup).
#include <type_traits>

#include "common.h"
#include "plan.h"

namespace dsht {

// Every segment's [first run, last run + 1); segments without runs keep (0, 0).
// Also clears nrmax per-set magnitude bounds.
__global__ __launch_bounds__(256) void k_seg_clear(int nseg, int* __restrict__ lo,
                                                   int* __restrict__ hi,
                                                   unsigned* __restrict__ rmax, int nrmax) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg) lo[s] = hi[s] = 0;
  if (s < nrmax) rmax[s] = 0u;
}

__global__ __launch_bounds__(256) void k_seg_runs(const int* __restrict__ run_seg,
                                                  const int* __restrict__ run_total,
                                                  int* __restrict__ lo, int* __restrict__ hi) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int nr = *run_total;
  if (r >= nr) return;
  const int s = run_seg[r];
  if (r == 0 || run_seg[r - 1] != s) lo[s] = r;
  if (r == nr - 1 || run_seg[r + 1] != s) hi[s] = r + 1;
}

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

constexpr int kFoldM = 64;   // orders per block (one wavefront)
constexpr int kFoldNp = 4;   // node pairs per block (one wavefront each)

template <typename TP, typename T, bool kY>
__global__ __launch_bounds__(kFoldM * kFoldNp) void k_fold(const TP* __restrict__ part,
                                                           const int* __restrict__ seg_lo,
                                                           const int* __restrict__ seg_hi,
                                                           int Nl, int Nx, int Npair, double dx,
                                                           int nsets, int set0,
                                                           int64_t part_stride, T* __restrict__ rhs,
                                                           unsigned* __restrict__ rmax) {
  using V = typename Vec2<TP>::type;
  // slot set0 + z of the batch, its partials part_stride elements further on
  const int set = set0 + (int)blockIdx.z;
  part += blockIdx.z * part_stride;
  if (rmax) rmax += blockIdx.z;
  const int m = blockIdx.y * kFoldM + (threadIdx.x & (kFoldM - 1));
  const int np = blockIdx.x * kFoldNp + threadIdx.x / kFoldM;
  const bool live = m < Nl && np < Npair;
  float vmax = 0.f;  // max |c|, |d / dx| written by this thread
  const V* pv = reinterpret_cast<const V*>(part);
  constexpr int kOut = kY ? 24 : 16;
  T out[2][kOut];  // [P][kOut]
  double dd[2][2][2][2] = {};  // d_P(n) of the Y-only form: [P][branch][re|im][node]
#pragma unroll
  for (int br = 0; br < 2 && live; ++br) {
    // c, d per (node of the pair, hemisphere), complex
    double c[2][2][2] = {}, d[2][2][2] = {};
    // the 8 (node, hemisphere, half) slots of this branch: their run ranges,
    // then the first run of every slot loaded at once (16 loads in flight),
    // then the sums in slot order (a segment's further runs, rare, after its
    // first)
    int r0[8], r1[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int odd = q >> 2, h = (q >> 1) & 1, half = q & 1;
      const int j = 2 * np + odd - half;  // bin n (v0, v1) or bin n-1 (v2, v3)
      r0[q] = r1[q] = 0;
      if (j >= 0 && j <= Nx - 2) {
        const int s = (br * 2 + h) * (Nx - 1) + j;
        r0[q] = seg_lo[s];
        r1[q] = seg_hi[s];
      }
    }
    V fc[8], fd[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int half = q & 1;
      const bool any = r0[q] < r1[q];
      const int64_t r = any ? r0[q] : 0;
      fc[q] = pv[(r * 4 + 2 * half) * Nl + m];
      fd[q] = pv[(r * 4 + 2 * half + 1) * Nl + m];
      if (!any) fc[q] = fd[q] = V{0, 0};
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int odd = q >> 2, h = (q >> 1) & 1, half = q & 1;
      c[odd][h][0] += (double)fc[q].x;
      c[odd][h][1] += (double)fc[q].y;
      d[odd][h][0] += (double)fd[q].x;
      d[odd][h][1] += (double)fd[q].y;
      for (int r = r0[q] + 1; r < r1[q]; ++r) {
        const V vc = pv[((int64_t)r * 4 + 2 * half) * Nl + m];
        const V vd = pv[((int64_t)r * 4 + 2 * half + 1) * Nl + m];
        c[odd][h][0] += (double)vc.x;
        c[odd][h][1] += (double)vc.y;
        d[odd][h][0] += (double)vd.x;
        d[odd][h][1] += (double)vd.y;
      }
    }
#pragma unroll
    for (int P = 0; P < 2; ++P) {
      const double sg = P ? -1.0 : 1.0;
#pragma unroll
      for (int ri = 0; ri < 2; ++ri)
#pragma unroll
        for (int odd = 0; odd < 2; ++odd) {
          const double cv = c[odd][0][ri] + sg * c[odd][1][ri];
          const double dv = d[odd][0][ri] + sg * d[odd][1][ri];
          if constexpr (kY) {
            out[P][2 * (2 * br + ri) + odd] = (T)cv;
            dd[P][br][ri][odd] = dx * dv;
          } else {
            out[P][8 * br + 4 * ri + odd] = (T)cv;
            out[P][8 * br + 4 * ri + 2 + odd] = (T)(dx * dv);
          }
          vmax = fmaxf(vmax, (float)fmax(fabs(cv), fabs(dv)));
        }
    }
  }
  if constexpr (kY) {
#pragma unroll
    for (int odd = 0; odd < 2; ++odd) {
      const int n = 2 * np + odd;
      const double x = (double)n * dx;  // node x_n
      const double inv = n < Nx ? 1.0 / (1.0 - x * x) : 0.0;  // padding nodes: d = 0
#pragma unroll
      for (int P = 0; P < 2; ++P)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          out[P][8 + 2 * q + odd] = (T)(x * inv * dd[P][q >> 1][q & 1][odd]);
          out[P][16 + 2 * q + odd] = (T)(inv * dd[1 - P][q >> 1][q & 1][odd]);
        }
    }
  }
  if (live) {
#pragma unroll
    for (int P = 0; P < 2; ++P) {
      if constexpr (kY) {
        // out[P][8 kind + 2 (2 br + ri) + node] -> branch br's record
#pragma unroll
        for (int br = 0; br < 2; ++br) {
          T* o = rhs + plain_rec(np, br, m, P, Nl);
#pragma unroll
          for (int q = 0; q < 12; q += 2) {
            const int kind = q >> 2, ri = (q >> 1) & 1;
            const int src = 8 * kind + 2 * (2 * br + ri);
            *reinterpret_cast<typename Vec2<T>::type*>(o + q) = {out[P][src], out[P][src + 1]};
          }
        }
      } else {
        T* o = rhs + ((((int64_t)m * 2 + P) * Npair + np) * nsets + set) * 16;
#pragma unroll
        for (int q = 0; q < kOut; q += 2)
          *reinterpret_cast<typename Vec2<T>::type*>(o + q) = {out[P][q], out[P][q + 1]};
      }
    }
  }
  // per-set magnitude bound for the matrix-core contraction's scaling (only
  // when that contraction follows): wavefront, then block, then one atomic
  if (rmax) {
    __shared__ float wmax[kFoldNp];
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = vmax;
    __syncthreads();
    if (threadIdx.x == 0) {
      float bm = wmax[0];
      for (int i = 1; i < kFoldNp; ++i) bm = fmaxf(bm, wmax[i]);
      atomic_max_pos(rmax, bm);
    }
  }
}

// All nsets slots of a weight-set batch (FP32, the 16-value layout).  A block
// takes 64 orders (lanes: every partial it reads is one 512-byte stretch) of
// one node pair, and the sets in chunks of kFoldSetChunk: each wavefront folds
// two sets of the chunk into an LDS tile, and the block then writes the
// chunk's rows rhs[m][P][np][chunk sets][16] as 512-byte stretches (rows of
// different sets are adjacent, and written all at once rather than a 64-byte
// piece at a time).  Per-set max(|c|, |d/dx|): wavefront, then one atomic per
// (block, set).
constexpr int kFoldSetChunk = 8;
constexpr int kFoldRow = kFoldSetChunk * 16 + 4;  // LDS row: 4 (mod 32) dwords of padding
// Weight sets share their runs (seg_stride 0, set s's partials part_stride
// further on); the point sets of a transform_many group have runs of their own
// (set s's run ranges seg_stride further on in seg_lo / seg_hi, part_stride
// 0).  The sets' rows go to slots set0 .. set0 + nsets - 1 of ntot.
__global__ __launch_bounds__(256) void k_fold_sets(
    const float* __restrict__ part, int64_t part_stride, const int* __restrict__ seg_lo,
    const int* __restrict__ seg_hi, int seg_stride, int Nl, int Nx, int Npair, double dx,
    int nsets, int set0, int ntot, float* __restrict__ rhs, unsigned* __restrict__ rmax) {
  __shared__ __attribute__((aligned(16))) float tile[2][kFoldM][kFoldRow];  // [P][m][set, 16]
  __shared__ float smx[kMaxSets];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m0 = blockIdx.y * kFoldM, m = m0 + lane;
  const int np = blockIdx.x;
  const bool live = m < Nl;
  // run ranges of the 8 (node, hemisphere, half) slots of each branch
  // (uniform: they depend on the node pair and the set's run arrays only)
  int r0[2][8], r1[2][8];
  auto slot_runs = [&](const int* lo, const int* hi) {
#pragma unroll
    for (int br = 0; br < 2; ++br)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int odd = q >> 2, h = (q >> 1) & 1, half = q & 1;
        const int j = 2 * np + odd - half;  // bin n (v0, v1) or bin n-1 (v2, v3)
        r0[br][q] = r1[br][q] = 0;
        if (j >= 0 && j <= Nx - 2) {
          const int sg = (br * 2 + h) * (Nx - 1) + j;
          r0[br][q] = lo[sg];
          r1[br][q] = hi[sg];
        }
      }
  };
  if (seg_stride == 0) slot_runs(seg_lo, seg_hi);
  for (int c0 = 0; c0 < nsets; c0 += kFoldSetChunk) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int sl = 2 * wv + u, set = c0 + sl;
      if (set >= nsets) break;  // uniform
      if (seg_stride != 0) slot_runs(seg_lo + set * seg_stride, seg_hi + set * seg_stride);
      const float2* pv = reinterpret_cast<const float2*>(part + set * part_stride);
      float vmax = 0.f;
      float out[2][16];
#pragma unroll
      for (int br = 0; br < 2; ++br) {
        double c[2][2][2] = {}, d[2][2][2] = {};
        float2 fc[8], fd[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int half = q & 1;
          const bool any = live && r0[br][q] < r1[br][q];
          const int64_t r = any ? r0[br][q] : 0;
          fc[q] = pv[(r * 4 + 2 * half) * Nl + (live ? m : 0)];
          fd[q] = pv[(r * 4 + 2 * half + 1) * Nl + (live ? m : 0)];
          if (!any) fc[q] = fd[q] = make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int odd = q >> 2, h = (q >> 1) & 1, half = q & 1;
          c[odd][h][0] += (double)fc[q].x;
          c[odd][h][1] += (double)fc[q].y;
          d[odd][h][0] += (double)fd[q].x;
          d[odd][h][1] += (double)fd[q].y;
          for (int r = r0[br][q] + 1; live && r < r1[br][q]; ++r) {
            const float2 vc = pv[((int64_t)r * 4 + 2 * half) * Nl + m];
            const float2 vd = pv[((int64_t)r * 4 + 2 * half + 1) * Nl + m];
            c[odd][h][0] += (double)vc.x;
            c[odd][h][1] += (double)vc.y;
            d[odd][h][0] += (double)vd.x;
            d[odd][h][1] += (double)vd.y;
          }
        }
#pragma unroll
        for (int P = 0; P < 2; ++P) {
          const double sg = P ? -1.0 : 1.0;
#pragma unroll
          for (int ri = 0; ri < 2; ++ri)
#pragma unroll
            for (int odd = 0; odd < 2; ++odd) {
              const double cv = c[odd][0][ri] + sg * c[odd][1][ri];
              const double dv = d[odd][0][ri] + sg * d[odd][1][ri];
              out[P][8 * br + 4 * ri + odd] = (float)cv;
              out[P][8 * br + 4 * ri + 2 + odd] = (float)(dx * dv);
              vmax = fmaxf(vmax, (float)fmax(fabs(cv), fabs(dv)));
            }
        }
      }
#pragma unroll
      for (int P = 0; P < 2; ++P)
#pragma unroll
        for (int q = 0; q < 16; q += 4)
          *reinterpret_cast<float4*>(&tile[P][lane][16 * sl + q]) =
              make_float4(out[P][q], out[P][q + 1], out[P][q + 2], out[P][q + 3]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
      if (lane == 0) smx[set] = vmax;
    }
    __syncthreads();
    // the chunk's rows: (m, P) -> nc sets x 16 floats, contiguous in rhs
    const int nc = min(kFoldSetChunk, nsets - c0);
    const int per_row = nc * 4;  // float4s per row
    for (int f = tid; f < 2 * kFoldM * per_row; f += 256) {
      const int row = f / per_row, within = f - row * per_row;
      const int ml = row >> 1, P = row & 1;
      if (m0 + ml < Nl)
        *reinterpret_cast<float4*>(
            rhs + ((((int64_t)(m0 + ml) * 2 + P) * Npair + np) * ntot + set0 + c0) * 16 +
            4 * within) = *reinterpret_cast<const float4*>(&tile[P][ml][4 * within]);
    }
    __syncthreads();
  }
  if (tid < nsets) atomic_max_pos(rmax + tid, smx[tid]);
}

static int fold_impl(Plan& pl, int nsets, int set0, int count, hipStream_t s, const void* part,
                     int64_t part_stride) {
  Workspace& ws = pl.ws;
  const int* run_total = ws.run_base + ws.cur_chunks;  // scan total = number of runs
  const bool scaled = use_mfma_contract(pl, nsets);
  const bool yform = !scaled && !pl.caller_tables;  // plain contraction of Y-only tables
  unsigned* rmax = scaled ? ws.rmax + set0 : nullptr;
  const int nclear = std::max(pl.nseg, scaled ? count : 0);
  k_seg_clear<<<(nclear + 255) / 256, 256, 0, s>>>(pl.nseg, ws.seg_run_lo, ws.seg_run_hi, rmax,
                                                   scaled ? count : 0);
  DSHT_LAUNCH_CHECK();
  const int64_t max_runs = (int64_t)ws.cur_chunks + pl.nseg + 1;
  k_seg_runs<<<(int)((max_runs + 255) / 256), 256, 0, s>>>(ws.run_seg, run_total, ws.seg_run_lo,
                                                          ws.seg_run_hi);
  DSHT_LAUNCH_CHECK();
  dim3 grid((pl.Npair + kFoldNp - 1) / kFoldNp, (pl.Nl + kFoldM - 1) / kFoldM, count);
  auto go = [&](auto tp, auto yf) {
    using T = decltype(tp);
    k_fold<T, T, decltype(yf)::value><<<grid, kFoldM * kFoldNp, 0, s>>>(
        (const T*)(part ? part : ws.part), ws.seg_run_lo, ws.seg_run_hi, pl.Nl, pl.Nx, pl.Npair,
        pl.dx, nsets, set0, part_stride, (T*)ws.rhs, rmax);
  };
  if (pl.precision == DSHT_FP32) {
    if (yform) go(0.f, std::true_type{});
    else go(0.f, std::false_type{});
  } else {
    if (yform) go(0.0, std::true_type{});
    else go(0.0, std::false_type{});
  }
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_fold(Plan& pl, int nsets, int set, hipStream_t s, const void* part) {
  return fold_impl(pl, nsets, set, 1, s, part, 0);
}

int launch_fold_sets(Plan& pl, int nsets, hipStream_t s, const void* part, int64_t part_stride,
                     int set0, int count) {
  Workspace& ws = pl.ws;
  if (count < 0) count = nsets;
  if (!use_mfma_contract(pl, nsets) || nsets > kMaxSets || set0 < 0 || count < 1 ||
      set0 + count > nsets) {
    set_error("dsht: the batched fold serves the matrix-core contraction (4..64 sets)");
    return DSHT_EINVAL;
  }
  const int* run_total = ws.run_base + ws.cur_chunks;  // scan total = number of runs
  k_seg_clear<<<(std::max(pl.nseg, count) + 255) / 256, 256, 0, s>>>(
      pl.nseg, ws.seg_run_lo, ws.seg_run_hi, ws.rmax + set0, count);
  DSHT_LAUNCH_CHECK();
  const int64_t max_runs = (int64_t)ws.cur_chunks + pl.nseg + 1;
  k_seg_runs<<<(int)((max_runs + 255) / 256), 256, 0, s>>>(ws.run_seg, run_total, ws.seg_run_lo,
                                                          ws.seg_run_hi);
  DSHT_LAUNCH_CHECK();
  dim3 grid(pl.Npair, (pl.Nl + kFoldM - 1) / kFoldM);
  k_fold_sets<<<grid, 256, 0, s>>>((const float*)part, part_stride, ws.seg_run_lo,
                                   ws.seg_run_hi, 0, pl.Nl, pl.Nx, pl.Npair, pl.dx, count, set0,
                                   nsets, (float*)ws.rhs, ws.rmax + set0);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_fold_many(Plan& pl, const ManySets& ms, int set0, int nsets, hipStream_t s) {
  Workspace& ws = pl.ws;
  if (!use_mfma_contract(pl, nsets) || set0 + ms.nsets > nsets || ms.nsets > ws.cap_segsets) {
    set_error("dsht: the point-set group fold serves the matrix-core contraction");
    return DSHT_EINVAL;
  }
  const int stride = pl.nseg + 2;
  const int nchunks = ms.chunk_off[ms.nsets];
  const int* run_total = ws.run_base + nchunks;  // scan total = number of runs
  const int nkeys = ms.nsets * stride;
  k_seg_clear<<<(std::max(nkeys, ms.nsets) + 255) / 256, 256, 0, s>>>(
      nkeys, ws.seg_run_lo, ws.seg_run_hi, ws.rmax + set0, ms.nsets);
  DSHT_LAUNCH_CHECK();
  const int64_t max_runs = (int64_t)nchunks + nkeys;
  k_seg_runs<<<(int)((max_runs + 255) / 256), 256, 0, s>>>(ws.run_seg, run_total, ws.seg_run_lo,
                                                          ws.seg_run_hi);
  DSHT_LAUNCH_CHECK();
  dim3 grid(pl.Npair, (pl.Nl + kFoldM - 1) / kFoldM);
  k_fold_sets<<<grid, 256, 0, s>>>((const float*)ws.part, 0, ws.seg_run_lo, ws.seg_run_hi, stride,
                                   pl.Nl, pl.Nx, pl.Npair, pl.dx, ms.nsets, set0, nsets,
                                   (float*)ws.rhs, ws.rmax + set0);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
