// Stage (c) for many right-hand sides at once (DSHT_FP32 plans): the table
// contraction of a batched transform as a matrix-core GEMM,
//   a_s(l, m) = sum_n Y_lm(x_n) c_s(m, n) + Y'_lm(x_n) d_s(m, n)
// for every set s (weight set or point set), both branches and both parts --
// the dot products of sht/interp_funcs_py.py:46-60, once per set in the
// reference, here with every table entry read once for all sets.
//
// Per row group (up to 64 rows l of one (m, P), the layout the plain
// contraction streams) the GEMM is
//   C[l][col] = sum_K T[l][K] R[K][col],
//   K = (node pair np, {Y(2np), Y(2np+1), Y'(2np), Y'(2np+1)}),
//   col = (set, branch, re|im)  (4 per set),
// run as v_mfma_f32_32x32x16_f16 with both operands split into fp16 hi + lo
// and three products hi*hi + hi*lo + lo*hi accumulated in FP32 (~2^-22 per
// product).  fp16 range: the Y' entries are multiplied by dx (|Y' dx| stays
// below ~12 for l < 2048) and the d coefficients divided by it; every set's
// coefficients are scaled by a power of two that brings max(|c|, |d / dx|) of
// the set (recorded by the fold) to [2^13, 2^14), and the result is scaled
// back exactly.
//
// One workgroup (8 wavefronts) per row group.  Per chunk of 16 node pairs
// (4 MFMA K steps) the workgroup stages the table block (64 rows x 64 K),
// split into hi/lo, in LDS in the MFMA A-operand layout (double-buffered, one
// barrier per chunk); wavefront w owns column tile w (32 columns = 8 sets)
// and both row tiles, loads its B operands straight from the node
// coefficients rhs[m][P][np][set][16] (32 consecutive columns = 512
// contiguous bytes per node pair) and issues 6 MFMAs per K step.  The next
// chunk's table entries and B operands are loaded into registers while the
// current chunk is multiplied.  Accumulators restart every chunk and are
// added into an outer FP32 sum.
#include "common.h"
#include "plan.h"
#include "split.h"

namespace dsht {

constexpr int kCK = 16;         // node pairs per chunk (4 MFMA K steps of 16)
constexpr int kCThreads = 512;  // 8 wavefronts

__global__ __launch_bounds__(kCThreads, 2) void k_contract_mfma(
    const float4* __restrict__ tab, const float* __restrict__ rhs,
    const GroupDesc* __restrict__ groups, int Npair, int Nl, int nsets, float dx,
    const unsigned* __restrict__ rmax, double* __restrict__ a_equ, double* __restrict__ a_pol) {
  const GroupDesc gd = groups[blockIdx.x];
  if (gd.nrows == 0) return;  // padding of the XCD schedule
  // [buffer][K step][row tile][hi|lo][lane]: 8 halves per lane, one 16-byte read
  __shared__ __attribute__((aligned(16))) u32x4 sA[2][4][2][2][64];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, h = l >> 5;
  const int ncol = 4 * nsets, nct = (ncol + 31) >> 5;
  const bool active = w < nct;  // wave-uniform
  const int col = 32 * w + (l & 31);
  const bool colv = active && col < ncol;
  const int set = colv ? col >> 2 : 0;
  const int br = (col >> 1) & 1, ri = col & 1;
  float bsc = 1.f;
  if (colv) {
    const float mx = __uint_as_float(rmax[set]);
    if (mx > 0.f) bsc = ldexpf(1.f, 13 - ilogbf(mx));
  }
  const float bscd = bsc / dx;
  const int64_t qstride = 16 * (int64_t)nsets;  // floats per node pair
  const float* qb = rhs + ((int64_t)(gd.m * 2 + gd.P) * Npair * nsets + set) * 16 + br * 8 + ri * 4;
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  const floatx16 zero = {};
  floatx16 out[2] = {zero, zero}, acc[2];

  // loads of one chunk: this thread's two table entries (node pair c0 + e/64,
  // row e%64, e = tid + 512 v) and this lane's B operands (K step t, node
  // pairs c0 + 4t + 2h + {0, 1})
  float4 ya[2], bq[4][2];
  // every load is issued unconditionally at a clamped index; out-of-range
  // entries are zeroed where the chunk is used, not where it is loaded, so
  // that the compiler waits for the loads only a chunk later (a select right
  // after a load waits for it at once; a conditional load of a stack zero
  // becomes a flat load, which LDS waits also wait for)
  typedef float v4 __attribute__((ext_vector_type(4)));
  const int rA = tid & 63, npA = 2 * (tid >> 6);  // this thread's table entries
  auto load = [&](int c0) {
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      // node pairs 2 (tid / 64) + v of the chunk: the two halves of one
      // A-operand element, so the staging writes whole 16-byte elements
      const int npc = min(c0 + npA + v, Npair - 1), rc = min(rA, gd.nrows - 1);
      const v4 t4 = __builtin_nontemporal_load(  // read once: non-temporal
          reinterpret_cast<const v4*>(tab + gd.off + (int64_t)npc * gd.nrows + rc));
      ya[v] = make_float4(t4.x, t4.y, t4.z, t4.w);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int np = c0 + 4 * t + 2 * h + u;
        bq[t][u] = *reinterpret_cast<const float4*>(qb + min(np, Npair - 1) * qstride);
      }
  };
  load(0);
  int buf = 0;
  for (int c0 = 0; c0 < Npair; c0 += kCK, buf ^= 1) {
    // stage this chunk's table block (split into hi/lo) into LDS buffer `buf`
    {
      // node pairs u = 2 (tid / 64) + half, half = 0, 1; K index within the
      // step: 4 (u % 4) + q -> lane half (u % 4) / 2, slot 4 ((u % 4) % 2) + q
      const int w = tid >> 6, r = tid & 63;
      const int t = w >> 1, hh = w & 1;
      const int rt = r >> 5, lane = hh * 32 + (r & 31);
      u32x4 hv, lv;
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        float4 y = c0 + npA + half < Npair && rA < gd.nrows ? ya[half] : zero4;
        y.z *= dx;
        y.w *= dx;
        uint32_t h01, l01, h23, l23;
        split2_f16(y.x, y.y, h01, l01);
        split2_f16(y.z, y.w, h23, l23);
        hv[2 * half] = h01;
        hv[2 * half + 1] = h23;
        lv[2 * half] = l01;
        lv[2 * half + 1] = l23;
      }
      sA[buf][t][rt][0][lane] = hv;
      sA[buf][t][rt][1][lane] = lv;
    }
    // B of this chunk, scaled; then the next chunk's loads go in flight
    float x[4][8];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const bool ok = colv && c0 + 4 * t + 2 * h + u < Npair;
        const float s0 = ok ? bsc : 0.f, s1 = ok ? bscd : 0.f;
        x[t][4 * u + 0] = bq[t][u].x * s0;
        x[t][4 * u + 1] = bq[t][u].y * s0;
        x[t][4 * u + 2] = bq[t][u].z * s1;
        x[t][4 * u + 3] = bq[t][u].w * s1;
      }
    if (c0 + kCK < Npair) load(c0 + kCK);
    // one barrier per chunk: every wave has finished the chunk that last
    // used buffer `buf` before anyone stages into it again
    __syncthreads();
    if (active) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        half8 bhi, blo;
        split8_f16(x[t], bhi, blo);
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
          const half8 ahi = __builtin_bit_cast(half8, sA[buf][t][rt][0][l]);
          const half8 alo = __builtin_bit_cast(half8, sA[buf][t][rt][1][l]);
          acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bhi, t == 0 ? zero : acc[rt], 0, 0, 0);
          acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, blo, acc[rt], 0, 0, 0);
          acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bhi, acc[rt], 0, 0, 0);
        }
      }
      out[0] += acc[0];
      out[1] += acc[1];
    }
  }

  if (colv) {
    // C layout: lane l, register r holds C[8 (r/4) + 4 (l/32) + r%4][l%32]
    const double inv = 1.0 / (double)bsc;  // exact: a power of two
    double* dst = br ? a_pol : a_equ;
    const int64_t base = (int64_t)set * ((int64_t)Nl * (Nl + 1) / 2);
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 32 * rt + 8 * (r >> 2) + 4 * h + (r & 3);
        if (row < gd.nrows) {
          const int ell = gd.m + gd.P + 2 * (gd.i0 + row);
          dst[(base + hp_index(ell, gd.m, Nl)) * 2 + ri] = (double)out[rt][r] * inv;
        }
      }
  }
}

int launch_contract_mfma(Plan& pl, const float* full32, int nsets, hipStream_t s) {
  if (4 * nsets > 32 * (kCThreads / 64)) {
    set_error("dsht: at most 64 sets per matrix-core contraction");
    return DSHT_EINVAL;
  }
  k_contract_mfma<<<pl.ngroups, kCThreads, 0, s>>>(
      reinterpret_cast<const float4*>(full32), reinterpret_cast<const float*>(pl.ws.rhs),
      pl.groups, pl.Npair, pl.Nl, nsets, (float)pl.dx, pl.ws.rmax, pl.ws.a_equ, pl.ws.a_pol);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
