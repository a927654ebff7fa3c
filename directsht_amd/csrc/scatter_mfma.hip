// Stage (b), part 2, matrix-core variant of the scatter (DSHT_FP32 plans):
//   v_k(m, seg) = sum_{p in seg} h_k(p) e^{-i m phi_p},  k = 0..3,
// the v's of sht/shared_interp_funcs.py:88-105, as a GEMM on the MFMA units.
//
// Factor m = T mt + 32 m1 + m2 (mt = blockIdx.y, m1 < M1, m2 < 32, m tile
// T = 32 M1 = 1024, 512 or 256 orders, the narrowest that covers Nl):
//   e^{-i m phi} = E(m1) B(m2),  E = e^{-i (T mt + 32 m1) phi},  B = e^{-i m2 phi}
// so for one run (stretch of one segment) of points p
//   V_k(m) = sum_p [h_k(p) E_p(m1)] B_p(m2)
// is a complex GEMM with rows (k, m1) and columns m2; written as one real GEMM
// with rows (k, m1, re|im) = 8 M1 and columns (re|im, m2) = 64:
//   V_re = C[re][re] - C[im][im],  V_im = C[re][im] + C[im][re].
// It runs as v_mfma_f32_32x32x16_f16 with each operand split into fp16
// hi + lo (x = hi + lo to ~2^-22 relative, or 2^-25 absolute where lo is
// subnormal, which the MFMA honours) and three products hi*hi + hi*lo + lo*hi
// accumulated in FP32 (the accumulator carries a whole run of at most one
// chunk of points; kGroup > 0 restarts it every kGroup windows into an outer
// FP32 sum).
//
// One workgroup (M1/4 wavefronts, one per 32-row tile) per (chunk of sorted
// points, m tile).  The chunk is staged KB points at a time (128 for the
// 1024-wide tile, 64 otherwise): thread (point p, group g) generates B for a
// range of m2 and E for a range of m1 by FP64 chains (steps e1 = e^{-i phi}
// and e1^32) into LDS, B split into fp16 hi/lo in [m2][point] order so every
// MFMA operand is one aligned 16-byte LDS read.  The block is then consumed in
// 16-aligned windows (one MFMA K step).  A window inside one run (almost all
// of them) builds A = h_k E unmasked; a window that contains a segment
// boundary is issued once per run, with the points of other runs masked to
// zero in A.  Wave w owns row tile w (row blocks 2w, 2w+1 of 8 values of m1
// each) and issues 6 MFMAs per window.  A run's result goes to its own slot
// part[run][k][m][2] (FP32), as for the vector scatter; the 32 lanes of a
// half-wave write 32 consecutive m (256 contiguous bytes).
//
// Operand layouts of v_mfma_f32_32x32x16_f16 (checked on gfx950):
//   A: lane l holds A[l%32][8*(l/32) + j],  j = 0..7
//   B: lane l holds B[8*(l/32) + j][l%32]
//   C: lane l, register r holds C[8*(r/4) + 4*(l/32) + r%4][l%32]
// Rows of a 32-row tile are ordered i = 16g + 8c + e (g: k parity, c: re|im,
// e: m1 offset), so the re and im rows of one (k, m1) sit in registers r and
// r + 4 of the same lane.
#include "common.h"
#include "plan.h"
#include "split.h"

namespace dsht {

namespace {

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

}  // namespace

constexpr int kWin = 16;    // points per MFMA K step
// Windows per MFMA accumulator restart into an outer FP32 sum; 0 = the MFMA
// accumulator carries the whole run.
constexpr int kGroup = 0;
// LDS rows are padded to 4 (mod 32) dwords (KB + 8 halves, KB + 4 floats):
// the 32 lanes of a half-wave read 32 different rows (m2 or m1) at the same
// column, and with this stride each group of 8 lanes covers all 32 banks with
// its 16-byte reads.

// A operand of one window pass: this lane's 8 points (K slice) of row
// (k, m1, re|im): h_k(p) E(m1, p), with points outside [p0, p1) zeroed when
// the window holds more than one run.
template <bool kMasked>
__device__ inline void build_a(const float (&hv)[8], const float (&ev)[8], int pg0, int p0, int p1,
                               half8& ahi, half8& alo) {
  float x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const float v = hv[j] * ev[j];
    x[j] = (!kMasked || (pg0 + j >= p0 && pg0 + j < p1)) ? v : 0.f;
  }
  split8_f16(x, ahi, alo);
}

// 6 MFMAs of one window pass: acc[c] (+)= A_hi B_hi + A_hi B_lo + A_lo B_hi
// for column tiles c = re, im.  kFresh restarts the accumulators from zero.
template <bool kFresh>
__device__ inline void mfma6(const half8& ahi, const half8& alo, const half8& brh, const half8& brl,
                             const half8& bih, const half8& bil, floatx16 (&acc)[2]) {
  const floatx16 zero = {};
  acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, brh, kFresh ? zero : acc[0], 0, 0, 0);
  acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, brl, acc[0], 0, 0, 0);
  acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, brh, acc[0], 0, 0, 0);
  acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bih, kFresh ? zero : acc[1], 0, 0, 0);
  acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bil, acc[1], 0, 0, 0);
  acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bih, acc[1], 0, 0, 0);
}

__device__ inline cplx cpow2k(cplx x, int k) {  // x^(2^k)
  for (int i = 0; i < k; ++i) x = csq(x);
  return x;
}

template <int M1, int KB>
__global__ __launch_bounds__(16 * M1, 4) void k_scatter_mfma(const PointRec* __restrict__ rec,
                                                          const int* __restrict__ seg_start,
                                                          int nseg, int chunk, int Nl,
                                                          const int* __restrict__ run_base,
                                                          const double* __restrict__ wsort,
                                                          float* __restrict__ part,
                                                          int* __restrict__ run_seg) {
  constexpr int kWaves = M1 / 4;              // row tiles: 4 k * M1 m1 * (re, im) / 32
  constexpr int kThreads = 64 * kWaves;
  constexpr int kBlk = KB;                    // points staged per generation phase
  constexpr int kRB = KB + 8;                 // padded row strides (4 mod 32 dwords)
  constexpr int kRE = KB + 4;
  constexpr int kGen = kThreads / kBlk;       // generation groups per point: 1, 2 or 4
  constexpr int kBPer = 32 / kGen;            // m2 values per generation thread
  constexpr int kEPer = M1 / kGen;            // m1 values per generation thread
  constexpr int kHPer = 4 / kGen;             // Hermite weights per generation thread
  constexpr int kQ = kBPer > kEPer ? kBPer : kEPer;
  constexpr int kMT = 32 * M1;                // orders per m tile
  constexpr int kMBlocks = M1 / 8;            // blocks of 8 m1 values per k
  static_assert(M1 % 8 == 0 && kThreads % kBlk == 0 && kGen <= 4 && kEPer >= 1,
                "unsupported m tile / staging block");
  const int c = blockIdx.x, mt = blockIdx.y;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;

  // B = e^{-i m2 phi} as fp16 hi/lo, [re|im][hi|lo][m2][point]; E FP32
  // [re|im][m1][point]; h_k FP32 [k][point]; segment of each point
  __shared__ __attribute__((aligned(16))) _Float16 sB[2][2][32][kRB];
  __shared__ __attribute__((aligned(16))) float sE[2][M1][kRE];
  __shared__ __attribute__((aligned(16))) float sH[4][kBlk];
  __shared__ int sSeg[kBlk];

  // wave w owns row tile w; its rows i = 16g + 8c + e are row block 2w + g
  // (k = (2w + g) / kMBlocks, m1 = 8 ((2w + g) % kMBlocks) + e), part c
  const int ia = l & 31, bq = l >> 5;
  const int gA = ia >> 4, cA = (ia >> 3) & 1;
  const int bA = 2 * w + gA;
  const int kA = bA / kMBlocks;
  const int m1A = 8 * (bA % kMBlocks) + (ia & 7);

  floatx16 acc[2], out[2];  // [column tile re | im]
  const floatx16 zero = {};
  acc[0] = acc[1] = out[0] = out[1] = zero;
  int wgroup = 0;

  // run tracking is wave-uniform (scalar registers, scalar branches)
  int run = __builtin_amdgcn_readfirstlane(run_base[c]);
  int cur = __builtin_amdgcn_readfirstlane(rec[b].seg);
  int run_end = min(e, seg_start[cur + 1]);

  for (int blk = b; blk < e; blk += kBlk) {
    const int bn = min(kBlk, e - blk);
    {
      // generation: thread (point p, group g, uniform per wavefront): B for
      // m2 in [kBPer g, kBPer (g+1)) and E for m1 in [kEPer g, kEPer (g+1))
      // by FP64 chains
      const int p = tid & (kBlk - 1);
      const int g = __builtin_amdgcn_readfirstlane(tid / kBlk);
      cplx e1{1.0, 0.0};
      int sg = -1;
      if (p < bn) {
        const PointRec& pr = rec[blk + p];
        e1 = cplx{pr.ere, pr.eim};
        const double ws = wsort ? wsort[blk + p] : 1.0;
        const double hk[4] = {pr.h0, pr.h1, pr.h2, pr.h3};
#pragma unroll
        for (int j = 0; j < kHPer; ++j) sH[g * kHPer + j][p] = (float)(ws * hk[g * kHPer + j]);
        sg = pr.seg;
      } else {
#pragma unroll
        for (int j = 0; j < kHPer; ++j) sH[g * kHPer + j][p] = 0.f;
      }
      if (g == 0) sSeg[p] = sg;
      const cplx e32 = cpow2k(e1, 5);
      constexpr int lgB = kBPer == 8 ? 3 : kBPer == 16 ? 4 : 5;
      constexpr int lgE = kEPer == 1 ? 0 : kEPer == 2 ? 1 : kEPer == 4 ? 2 : kEPer == 8 ? 3 : 4;
      constexpr int lgM1 = M1 == 8 ? 3 : M1 == 16 ? 4 : 5;
      cplx bz = cpow(cpow2k(e1, lgB), g);             // e1^(kBPer g)
      cplx ez = cpow(cpow2k(e32, lgE), g);            // e1^(32 kEPer g)
      if (mt > 0) ez = cmul(ez, cpow(cpow2k(e32, lgM1), mt));  // * e1^(kMT mt)
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        if (q < kBPer) {
          const int m = kBPer * g + q;
          uint32_t hi, lo;
          split2_f16((float)bz.x, (float)bz.y, hi, lo);
          sB[0][0][m][p] = half_lo(hi);
          sB[1][0][m][p] = half_hi(hi);
          sB[0][1][m][p] = half_lo(lo);
          sB[1][1][m][p] = half_hi(lo);
          bz = cmul(bz, e1);
        }
        if (q < kEPer) {
          sE[0][kEPer * g + q][p] = (float)ez.x;
          sE[1][kEPer * g + q][p] = (float)ez.y;
          ez = cmul(ez, e32);
        }
      }
    }
    __syncthreads();
    for (int wo = 0; wo < bn; wo += kWin) {
      const int whi = blk + min(wo + kWin, bn);
      int p = blk + wo;
      // operands of this window that do not depend on the run split
      const int po = wo + 8 * bq;  // first point of this lane's K slice
      const float4 E0 = *reinterpret_cast<const float4*>(&sE[cA][m1A][po]);
      const float4 E1 = *reinterpret_cast<const float4*>(&sE[cA][m1A][po + 4]);
      const float4 H0 = *reinterpret_cast<const float4*>(&sH[kA][po]);
      const float4 H1 = *reinterpret_cast<const float4*>(&sH[kA][po + 4]);
      const half8 brh = *reinterpret_cast<const half8*>(&sB[0][0][l & 31][po]);
      const half8 brl = *reinterpret_cast<const half8*>(&sB[0][1][l & 31][po]);
      const half8 bih = *reinterpret_cast<const half8*>(&sB[1][0][l & 31][po]);
      const half8 bil = *reinterpret_cast<const half8*>(&sB[1][1][l & 31][po]);
      const float ev[8] = {E0.x, E0.y, E0.z, E0.w, E1.x, E1.y, E1.z, E1.w};
      const float hv[8] = {H0.x, H0.y, H0.z, H0.w, H1.x, H1.y, H1.z, H1.w};
      while (p < whi) {
        const int vhi = min(whi, run_end);  // points [p, vhi) belong to run `cur`
        half8 ahi, alo;
        const int pg0 = blk + po;
        if (vhi == whi && p == blk + wo)  // uniform: the whole window is one run
          build_a<false>(hv, ev, pg0, p, vhi, ahi, alo);
        else
          build_a<true>(hv, ev, pg0, p, vhi, ahi, alo);
        if (wgroup == 0)  // uniform
          mfma6<true>(ahi, alo, brh, brl, bih, bil, acc);
        else
          mfma6<false>(ahi, alo, brh, brl, bih, bil, acc);
        p = vhi;
        const bool run_done = p == run_end;
        if constexpr (kGroup > 0) {
          if (++wgroup == kGroup || run_done) {
            out[0] += acc[0];
            out[1] += acc[1];
            wgroup = 0;
          }
        } else {
          wgroup = run_done ? 0 : 1;
        }
        if (run_done) {
          // flush: registers 8g + d (re row) and 8g + 4 + d (im row) hold row
          // block 2w + g: k = (2w+g) / kMBlocks, m1 = 8 ((2w+g) % kMBlocks) + 4(l/32) + d
#pragma unroll
          for (int g = 0; g < 2; ++g) {
            const int bb = 2 * w + g;
            const int kk = bb / kMBlocks;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
              const int m = kMT * mt + 32 * (8 * (bb % kMBlocks) + 4 * bq + d) + (l & 31);
              if (m < Nl) {
                const floatx16& r0 = kGroup > 0 ? out[0] : acc[0];
                const floatx16& r1 = kGroup > 0 ? out[1] : acc[1];
                const float vre = r0[8 * g + d] - r1[8 * g + 4 + d];
                const float vim = r1[8 * g + d] + r0[8 * g + 4 + d];
                *reinterpret_cast<float2*>(part + (((int64_t)run * 4 + kk) * Nl + m) * 2) =
                    make_float2(vre, vim);
              }
            }
          }
          if (tid == 0 && mt == 0) run_seg[run] = cur;
          if constexpr (kGroup > 0) out[0] = out[1] = zero;
          if (p < e) {
            ++run;
            const int off = p - blk;  // point p may lie past this block
            if (off < bn)
              cur = __builtin_amdgcn_readfirstlane(sSeg[off]);
            else
              cur = __builtin_amdgcn_readfirstlane(rec[p].seg);
            run_end = min(e, seg_start[cur + 1]);
          }
        }
      }
    }
    __syncthreads();
  }
}

template <int M1, int KB>
static int launch_m1(Plan& pl, int nchunks, const double* wsort, hipStream_t s) {
  dim3 grid(nchunks, (pl.Nl + 32 * M1 - 1) / (32 * M1));
  k_scatter_mfma<M1, KB><<<grid, 16 * M1, 0, s>>>(pl.ws.rec, pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                              pl.ws.run_base, wsort, (float*)pl.ws.part,
                                              pl.ws.run_seg);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

// m tile of 1024, 512 or 256 orders: the narrowest that covers Nl, or 1024
int launch_scatter_mfma(Plan& pl, int64_t Np, const double* wsort, hipStream_t s) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  if (pl.Nl <= 256) return launch_m1<8, 64>(pl, nchunks, wsort, s);
  if (pl.Nl <= 512) return launch_m1<16, 64>(pl, nchunks, wsort, s);
  return launch_m1<32, 128>(pl, nchunks, wsort, s);
}

}  // namespace dsht
