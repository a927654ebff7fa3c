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
#include <type_traits>

#include "common.h"
#include "plan.h"
#include "split.h"

namespace dsht {

constexpr int kCK = 16;         // node pairs per chunk (4 MFMA K steps of 16)
constexpr int kCThreads = 512;  // 8 wavefronts
constexpr int kCRT = 4;         // row tiles of 32: two table groups of up to 64 rows
static_assert(kGroupRows == 64, "a work item is two table groups of 64 rows");

__global__ __launch_bounds__(kCThreads, 1) void k_contract_mfma(
    const float4* __restrict__ tab, const float* __restrict__ rhs,
    const GroupDesc* __restrict__ groups, int Npair, int Nl, int nsets, float dx,
    const unsigned* __restrict__ rmax, double* __restrict__ a_equ, double* __restrict__ a_pol) {
  const GroupDesc gd = groups[blockIdx.x];
  if (gd.nrows == 0) return;  // padding of the XCD schedule
  // the pair's table groups: rows [0, na) at gd.off (row tiles 0, 1) and
  // rows [64, 64 + nb) at gd.off + 64 Npair (row tiles 2, 3)
  const int na = min(gd.nrows, 64), nb = gd.nrows - na;
  // [buffer][K step][row tile][hi|lo][lane]: 8 halves per lane, one 16-byte read
  __shared__ __attribute__((aligned(16))) u32x4 sA[2][4][kCRT][2][64];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, h = l >> 5;
  const int ncol = 4 * nsets, nct = (ncol + 31) >> 5;
  const bool active = w < nct;  // wave-uniform
  const int col = 32 * w + (l & 31);
  const bool colv = active && col < ncol;
  const int set = colv ? col >> 2 : 0;
  const int br = (col >> 1) & 1, ri = col & 1;
  float bsc = 0.f;  // padding columns: zero operands
  if (colv) {
    const float mx = __uint_as_float(rmax[set]);
    bsc = mx > 0.f ? ldexpf(1.f, 13 - ilogbf(mx)) : 1.f;
  }
  const float bscd = bsc / dx;
  // B: node pair np of the (m, P) slice at qs + np qstride (uniform), this
  // lane's column and node-pair half (2h) at qlane
  const int64_t qstride = 16 * (int64_t)nsets;  // floats per node pair
  const float* qs = rhs + (int64_t)(gd.m * 2 + gd.P) * Npair * qstride;
  const unsigned qlane = (unsigned)(2 * h * qstride + set * 16 + br * 8 + ri * 4);
  const floatx16 zero = {};
  floatx16 out[kCRT] = {zero, zero, zero, zero}, acc[kCRT];

  // loads of one chunk of node pairs [cs, cs + kCK): this thread's table
  // entries (row rA of each group, node pairs cs + npA + v) and this lane's B
  // operands (K step t, node pairs cs + 4t + 2h + {0, 1}), each from a
  // uniform base (scalar registers) and a fixed 32-bit lane offset.  The
  // last chunk is the window ending at Npair (cs = Npair - kCK, so no index
  // leaves the slice); its table entries below the chunk start are zeroed
  // when staged.
  float4 ya[2][2], bq[4][2];
  typedef float v4 __attribute__((ext_vector_type(4)));
  const int rA = tid & 63, npA = 2 * (tid >> 6);
  const int nrow[2] = {na, nb};
  unsigned toff[2][2], qoff[4][2];
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int v = 0; v < 2; ++v)
      toff[g][v] = (unsigned)(((npA + v) * nrow[g] + min(rA, max(nrow[g], 1) - 1)) * 16);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u) qoff[t][u] = (unsigned)(((4 * t + u) * qstride + qlane) * 4);
  const float4* tgrp[2] = {tab + gd.off, tab + gd.off + 64 * (int64_t)Npair};
  auto load = [&](int cs) {
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      if (g == 1 && nb == 0) break;  // uniform
      const char* tb = reinterpret_cast<const char*>(tgrp[g] + (int64_t)cs * nrow[g]);
#pragma unroll
      for (int v = 0; v < 2; ++v) {  // read once: non-temporal
        const v4 t4 = __builtin_nontemporal_load(reinterpret_cast<const v4*>(tb + toff[g][v]));
        ya[g][v] = make_float4(t4.x, t4.y, t4.z, t4.w);
      }
    }
    const char* qb = reinterpret_cast<const char*>(qs + (int64_t)cs * qstride);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int u = 0; u < 2; ++u) bq[t][u] = *reinterpret_cast<const float4*>(qb + qoff[t][u]);
  };
  // the main loop for NT row tiles (rows 0 .. 32 NT - 1 hold the rows: a
  // prefix, since group a fills before group b)
  auto run = [&](auto ntc) {
  constexpr int NT = decltype(ntc)::value;
  load(0);
  int buf = 0;
  for (int c0 = 0; c0 < Npair; c0 += kCK, buf ^= 1) {
    const int cs = min(c0, Npair - kCK);  // this chunk's window
    // stage this chunk's table block (split into hi/lo) into LDS buffer `buf`
    {
      // node pairs u = 2 (tid / 64) + half, half = 0, 1; K index within the
      // step: 4 (u % 4) + q -> lane half (u % 4) / 2, slot 4 ((u % 4) % 2) + q
      const int t = w >> 1, hh = w & 1;
      const int rt = rA >> 5, lane = hh * 32 + (rA & 31);
#pragma unroll
      for (int g = 0; g < (NT > 2 ? 2 : 1); ++g) {
        u32x4 hv, lv;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          float4 y = ya[g][half];
          if (cs + npA + half < c0) y = make_float4(0.f, 0.f, 0.f, 0.f);  // done before
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
        sA[buf][t][2 * g + rt][0][lane] = hv;
        sA[buf][t][2 * g + rt][1][lane] = lv;
      }
    }
    // B of this chunk, scaled; then the next chunk's loads go in flight
    float x[4][8];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        x[t][4 * u + 0] = bq[t][u].x * bsc;
        x[t][4 * u + 1] = bq[t][u].y * bsc;
        x[t][4 * u + 2] = bq[t][u].z * bscd;
        x[t][4 * u + 3] = bq[t][u].w * bscd;
      }
    if (c0 + kCK < Npair) load(min(c0 + kCK, Npair - kCK));
    // one barrier per chunk: every wave has finished the chunk that last
    // used buffer `buf` before anyone stages into it again
    __syncthreads();
    if (active) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        half8 bhi, blo;
        split8_f16(x[t], bhi, blo);
#pragma unroll
        for (int rt = 0; rt < NT; ++rt) {
          const half8 ahi = __builtin_bit_cast(half8, sA[buf][t][rt][0][l]);
          const half8 alo = __builtin_bit_cast(half8, sA[buf][t][rt][1][l]);
          acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bhi, t == 0 ? zero : acc[rt], 0, 0, 0);
          acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, blo, acc[rt], 0, 0, 0);
          acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bhi, acc[rt], 0, 0, 0);
        }
      }
#pragma unroll
      for (int rt = 0; rt < NT; ++rt) out[rt] += acc[rt];
    }
  }

  };
  switch ((gd.nrows + 31) / 32) {  // uniform
    case 1: run(std::integral_constant<int, 1>{}); break;
    case 2: run(std::integral_constant<int, 2>{}); break;
    case 3: run(std::integral_constant<int, 3>{}); break;
    default: run(std::integral_constant<int, 4>{}); break;
  }

  if (colv) {
    // C layout: lane l, register r holds C[8 (r/4) + 4 (l/32) + r%4][l%32]
    const double inv = 1.0 / (double)bsc;  // exact: a power of two
    double* dst = br ? a_pol : a_equ;
    const int64_t base = (int64_t)set * ((int64_t)Nl * (Nl + 1) / 2);
#pragma unroll
    for (int rt = 0; rt < kCRT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int rr = 32 * (rt & 1) + 8 * (r >> 2) + 4 * h + (r & 3);  // row within its group
        if (rr < (rt < 2 ? na : nb)) {
          const int row = 64 * (rt >> 1) + rr;
          const int ell = gd.m + gd.P + 2 * (gd.i0 + row);
          dst[(base + hp_index(ell, gd.m, Nl)) * 2 + ri] = (double)out[rt][r] * inv;
        }
      }
  }
}

int launch_contract_mfma(Plan& pl, const float* full32, int nsets, hipStream_t s) {
  if (4 * nsets > 32 * (kCThreads / 64) || pl.Npair < kCK) {
    set_error("dsht: the matrix-core contraction takes at most 64 sets and 16 node pairs");
    return DSHT_EINVAL;
  }
  k_contract_mfma<<<pl.nmgroups, kCThreads, 0, s>>>(
      reinterpret_cast<const float4*>(full32), reinterpret_cast<const float*>(pl.ws.rhs),
      pl.mgroups, pl.Npair, pl.Nl, nsets, (float)pl.dx, pl.ws.rmax, pl.ws.a_equ, pl.ws.a_pol);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
