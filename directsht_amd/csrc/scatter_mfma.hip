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
// chunk of points).
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
// each) and issues 6 MFMAs per window; the next window's LDS loads and A build
// are interleaved with them (a wavefront's VALU work overlaps its own MFMAs,
// not those of other wavefronts on the SIMD).  A run's result goes to its own slot
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
#include <type_traits>

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

// Empty register pins: a value pinned at a scheduling fence is computed on
// that side of it (no hoisting above, no sinking into later blocks).
template <typename T>
__device__ inline void pin(T& x) {
  asm volatile("" : "+v"(x));
}
__device__ inline void pin(cplx& z) {
  pin(z.x);
  pin(z.y);
}

// kGather: the weights are w[rec.idx] of a weight set in the caller's input
// order (wsort = w), gathered here; otherwise wsort[sorted position] or, if
// NULL, 1 (the records' own weights).
// The scatter of one work item: sorted points [b, e) of a set whose segment
// starts are seg_start, runs numbered from run0 (recorded in run_seg as
// key_off + segment), m tile mt.
template <int M1, int KB, bool kGather>
__device__ __forceinline__ void scatter_mfma_item(const PointRec* __restrict__ rec,
                                         const int* __restrict__ seg_start, const int b,
                                         const int e, const int run0, const int key_off,
                                         const int Nl, const int mt,
                                         const double* __restrict__ wsort,
                                         float* __restrict__ part, int* __restrict__ run_seg,
                                         const unsigned* __restrict__ wmax) {
  constexpr int kWaves = M1 / 4;              // row tiles: 4 k * M1 m1 * (re, im) / 32
  constexpr int kThreads = 64 * kWaves;
  constexpr int kBlk = KB;                    // points per staged block
  constexpr int kRB = KB + 8;                 // padded row strides (4 mod 32 dwords)
  constexpr int kRE = KB + 4;
  constexpr int kGen = kThreads / kBlk;       // generation groups per point: 1, 2 or 4
  constexpr int kBPer = 32 / kGen;            // m2 values per generation thread
  constexpr int kEPer = M1 / kGen;            // m1 values per generation thread
  constexpr int kHPer = 4 / kGen;             // Hermite weights per generation thread
  constexpr int kNWin = KB / kWin;            // windows per staged block
  constexpr int kBStep = kBPer / kNWin;       // chain steps per window
  constexpr int kEStep = kEPer / kNWin;
  constexpr int kMT = 32 * M1;                // orders per m tile
  constexpr int kMBlocks = M1 / 8;            // blocks of 8 m1 values per k
  static_assert(M1 % 8 == 0 && kThreads % kBlk == 0 && kGen <= 4 && kGen >= 1,
                "unsupported m tile / staging block");
  static_assert(kBStep * kNWin == kBPer && kEStep * kNWin == kEPer && kEStep >= 1,
                "chain steps must spread evenly over the windows of a block");
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;

  // two staging buffers: block i is multiplied out of one while block i+1 is
  // generated into the other.  B = e^{-i m2 phi} as fp16 hi/lo,
  // [re|im][hi|lo][m2][point]; E FP32 [re|im][m1][point]; h_k FP32
  // [k][point]; segment of each point
  __shared__ __attribute__((aligned(16))) _Float16 sB[2][2][2][32][kRB];
  __shared__ __attribute__((aligned(16))) float sE[2][2][M1][kRE];
  __shared__ __attribute__((aligned(16))) float sH[2][4][kBlk];
  __shared__ int sSeg[2][kBlk];

  // wave w owns row tile w; its rows i = 16g + 8c + e are row block 2w + g
  // (k = (2w + g) / kMBlocks, m1 = 8 ((2w + g) % kMBlocks) + e), part c
  const int ia = l & 31, bq = l >> 5;
  const int gA = ia >> 4, cA = (ia >> 3) & 1;
  const int bA = 2 * w + gA;
  const int kA = bA / kMBlocks;
  const int m1A = 8 * (bA % kMBlocks) + (ia & 7);

  // Range of the fp16 operands: the staged weights w h_k are scaled by
  // 2^-ex, with 2^ex the power of two at or above max |w| of the transform
  // (|h_k| <= 1 and |E| = 1, so every operand is at most 1 in magnitude and
  // the split keeps its ~2^-22 relative precision for any weight range), and
  // each run's sums are scaled back by 2^ex, exactly.
  int ex = 0;
  {
    const float mx = __uint_as_float(wmax[0]);
    if (mx > 0.f && mx <= 3.0e38f) (void)frexpf(mx, &ex);
    ex = max(-120, min(120, ex));
  }
  const double w_down = ldexp(1.0, -ex);
  const float w_up = ldexpf(1.0f, ex);

  floatx16 acc[2];  // [column tile re | im], accumulating the current run
  const floatx16 zero = {};
  acc[0] = acc[1] = zero;

  // run tracking is wave-uniform (scalar registers, scalar branches)
  int run = __builtin_amdgcn_readfirstlane(run0);
  int cur = __builtin_amdgcn_readfirstlane(rec[b].seg);
  int run_end = min(e, seg_start[cur + 1]);
  int blk = b, bn = 0, buf = 0;  // staged block: first point, size, buffer

  // ---- generation: thread (point gp, group gg) makes B for m2 = gg + kGen q
  // (q < kBPer) and E for m1 = gg + kGen q (q < kEPer) by FP64 chains with
  // steps e1^kGen and e1^(32 kGen), e1 = e^{-i phi}
  const int gp = tid % kBlk;
  const int gg = __builtin_amdgcn_readfirstlane(tid / kBlk);
  cplx gb{1.0, 0.0}, ge{1.0, 0.0}, gbs{1.0, 0.0}, ges{1.0, 0.0};  // chain values and steps
  // record of this thread's point in a later block, loaded one block ahead
  double pf_ere = 1.0, pf_eim = 0.0, pf_w = 0.0, pf_h[kHPer];
  int pf_seg = -1, pf_idx = 0;
  bool pf_ok = false;
#pragma unroll
  for (int j = 0; j < kHPer; ++j) pf_h[j] = 0.0;
  // h and weight of this thread's point in the block being generated; the
  // weight (gathered: loaded at gen_setup) is combined into sH at the end of
  // the current block's windows
  double hold_h[kHPer], hold_w = 0.0, hold_ok = 0.0;
  auto prefetch = [&](int pblk) {
    const int p = pblk + gp;
    if (p < e) {
      // loaded into registers and first used a block later: no value is
      // selected here (that would wait for the loads)
      const PointRec& pr = rec[p];
      pf_ere = pr.ere;
      pf_eim = pr.eim;
      const double* hk = &pr.h0 + gg * kHPer;  // h0..h3 are consecutive
#pragma unroll
      for (int j = 0; j < kHPer; ++j) pf_h[j] = hk[j];
      pf_seg = pr.seg;
      if constexpr (kGather) pf_idx = pr.idx;
      else pf_w = wsort ? wsort[p] : 1.0;
      pf_ok = true;
    } else {
      pf_ere = 1.0;
      pf_eim = 0.0;
      pf_w = 0.0;
      pf_seg = -1;
      pf_idx = 0;
      pf_ok = false;
    }
  };
  // chain start for the prefetched block, h and segments into buffer nbuf
  auto gen_setup = [&](int nbuf) {
#pragma unroll
    for (int j = 0; j < kHPer; ++j) hold_h[j] = pf_h[j];
    // unconditional load (index 0 for padding points, scaled by hold_ok = 0):
    // nothing waits for it until write_h
    if constexpr (kGather) hold_w = wsort[pf_idx];
    else hold_w = pf_w;
    hold_ok = pf_ok ? 1.0 : 0.0;
    if (gg == 0) sSeg[nbuf][gp] = pf_seg;
    const cplx e1{pf_ere, pf_eim};
    const cplx e2 = csq(e1), e4 = csq(e2), e32 = cpow2k(e4, 3);
    const cplx e64 = csq(e32), e128 = csq(e64);
    gbs = kGen == 4 ? e4 : kGen == 2 ? e2 : e1;                 // e1^kGen
    ges = kGen == 4 ? e128 : kGen == 2 ? e64 : e32;             // e1^(32 kGen)
    gb = gg == 0 ? cplx{1.0, 0.0} : gg == 1 ? e1 : gg == 2 ? e2 : cmul(e2, e1);        // e1^gg
    ge = gg == 0 ? cplx{1.0, 0.0} : gg == 1 ? e32 : gg == 2 ? e64 : cmul(e64, e32);    // e1^(32 gg)
    if (mt > 0) {
      constexpr int lgT = M1 == 8 ? 3 : M1 == 16 ? 4 : 5;     // kMT = 32 * 2^lgT
      ge = cmul(ge, cpow(cpow2k(e32, lgT), mt));               // * e1^(kMT mt)
    }
  };
  auto write_h = [&](int nbuf) {
#pragma unroll
    for (int j = 0; j < kHPer; ++j)
      sH[nbuf][gg * kHPer + j][gp] = (float)(hold_ok * hold_w * w_down * hold_h[j]);
  };
  // chain steps of window j of the block being generated into buffer nbuf
  auto gen_slice_b = [&](int j, int nbuf) {
    pin(gb);
#pragma unroll
    for (int t = 0; t < kBStep; ++t) {
      const int m2 = gg + kGen * (j * kBStep + t);
      uint32_t hi, lo;
      split2_f16((float)gb.x, (float)gb.y, hi, lo);
      sB[nbuf][0][0][m2][gp] = half_lo(hi);
      sB[nbuf][1][0][m2][gp] = half_hi(hi);
      sB[nbuf][0][1][m2][gp] = half_lo(lo);
      sB[nbuf][1][1][m2][gp] = half_hi(lo);
      gb = cmul(gb, gbs);
    }
    pin(gb);
  };
  auto gen_slice_e = [&](int j, int nbuf) {
    pin(ge);
#pragma unroll
    for (int t = 0; t < kEStep; ++t) {
      const int m1 = gg + kGen * (j * kEStep + t);
      sE[nbuf][0][m1][gp] = (float)ge.x;
      sE[nbuf][1][m1][gp] = (float)ge.y;
      ge = cmul(ge, ges);
    }
    pin(ge);
  };
  auto gen_slice = [&](int j, int nbuf) {
    gen_slice_b(j, nbuf);
    gen_slice_e(j, nbuf);
  };

  // ---- windows
  // this lane's E and h slices of window wo (8 points from wo + 8 (l/32))
  typedef float f32x4 __attribute__((ext_vector_type(4)));  // 16-byte aligned LDS reads
  auto load_eh = [&](int wo, float (&ev)[8], float (&hv)[8]) {
    const int po = wo + 8 * bq;
    const f32x4 E0 = *reinterpret_cast<const f32x4*>(&sE[buf][cA][m1A][po]);
    const f32x4 E1 = *reinterpret_cast<const f32x4*>(&sE[buf][cA][m1A][po + 4]);
    const f32x4 H0 = *reinterpret_cast<const f32x4*>(&sH[buf][kA][po]);
    const f32x4 H1 = *reinterpret_cast<const f32x4*>(&sH[buf][kA][po + 4]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ev[j] = E0[j];
      ev[4 + j] = E1[j];
      hv[j] = H0[j];
      hv[4 + j] = H1[j];
    }
  };
  // this lane's B operands of window wo: re hi, re lo, im hi, im lo
  auto load_b = [&](int wo, half8 (&bv)[4]) {
    const int po = wo + 8 * bq;
    bv[0] = *reinterpret_cast<const half8*>(&sB[buf][0][0][l & 31][po]);
    bv[1] = *reinterpret_cast<const half8*>(&sB[buf][0][1][l & 31][po]);
    bv[2] = *reinterpret_cast<const half8*>(&sB[buf][1][0][l & 31][po]);
    bv[3] = *reinterpret_cast<const half8*>(&sB[buf][1][1][l & 31][po]);
  };
  // the current run ends at point p: write its slot, clear the accumulators,
  // move to the next run
  auto finish_run = [&](int p) {
    // registers 8g + d (re row) and 8g + 4 + d (im row) hold row block
    // 2w + g: k = (2w+g) / kMBlocks, m1 = 8 ((2w+g) % kMBlocks) + 4(l/32) + d
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int bb = 2 * w + g;
      const int kk = bb / kMBlocks;
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const int m = kMT * mt + 32 * (8 * (bb % kMBlocks) + 4 * bq + d) + (l & 31);
        if (m < Nl) {
          const float vre = (acc[0][8 * g + d] - acc[1][8 * g + 4 + d]) * w_up;
          const float vim = (acc[1][8 * g + d] + acc[0][8 * g + 4 + d]) * w_up;
          *reinterpret_cast<float2*>(part + (((int64_t)run * 4 + kk) * Nl + m) * 2) =
              make_float2(vre, vim);
        }
      }
    }
    acc[0] = acc[1] = zero;
    if (tid == 0 && mt == 0) run_seg[run] = key_off + cur;
    if (p < e) {
      ++run;
      const int off = p - blk;  // point p may lie past this block
      if (off < bn)
        cur = __builtin_amdgcn_readfirstlane(sSeg[buf][off]);
      else
        cur = __builtin_amdgcn_readfirstlane(rec[p].seg);
      run_end = min(e, seg_start[cur + 1]);
    }
  };
  // window wo: MFMAs with its operands (cb, ca); interleaved with them, the
  // next window's operands (nb, na), if kNext, and, if kGenNext, slice
  // wo / kWin of the next block's generation
  auto step = [&](auto gen_next, auto next, int wo, half8 (&cb)[4], half8 (&ca)[2],
                  half8 (&nb)[4], half8 (&na)[2]) {
    constexpr bool kGenNext = decltype(gen_next)::value;
    constexpr bool kNext = decltype(next)::value;
    const int whi = blk + min(wo + kWin, bn);
    const int j = wo / kWin;
    float ev[8], hv[8];
    if (whi <= run_end) {  // uniform: the whole window lies in the current run
      // explicit order (scheduling fences): the LDS loads first, then each
      // MFMA followed by a share of the VALU work, so that the wavefront
      // issues that work while its MFMAs execute
      if constexpr (kNext) {
        load_eh(wo + kWin, ev, hv);
        load_b(wo + kWin, nb);
      }
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ca[0], cb[0], acc[0], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (kGenNext) gen_slice_b(j, buf ^ 1);
      __builtin_amdgcn_sched_barrier(0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ca[0], cb[2], acc[1], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (kGenNext) gen_slice_e(j, buf ^ 1);
      __builtin_amdgcn_sched_barrier(0);
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ca[0], cb[1], acc[0], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      uint32_t nh[4], nl[4];
      if constexpr (kNext) {  // next A, hi parts: products and their fp16 rounding
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float2v v = {hv[2 * q] * ev[2 * q], hv[2 * q + 1] * ev[2 * q + 1]};
          nh[q] = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, half2v));
          pin(nh[q]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ca[0], cb[3], acc[1], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (kNext) {  // next A, lo parts: fp16(x - hi), x - hi exact in FP32
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const half2v hh = __builtin_bit_cast(half2v, nh[q]);
          half2v ll;
          ll[0] = (_Float16)__builtin_fmaf((float)hh[0], -1.0f, hv[2 * q] * ev[2 * q]);
          ll[1] = (_Float16)__builtin_fmaf((float)hh[1], -1.0f, hv[2 * q + 1] * ev[2 * q + 1]);
          nl[q] = __builtin_bit_cast(uint32_t, ll);
          pin(nl[q]);
        }
        na[0] = __builtin_bit_cast(half8, (u32x4){nh[0], nh[1], nh[2], nh[3]});
        na[1] = __builtin_bit_cast(half8, (u32x4){nl[0], nl[1], nl[2], nl[3]});
      }
      __builtin_amdgcn_sched_barrier(0);
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ca[1], cb[0], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ca[1], cb[2], acc[1], 0, 0, 0);
      if (whi == run_end) finish_run(whi);
    } else {
      // a segment boundary inside the window: one masked pass per run
      load_eh(wo, ev, hv);
      const int pg0 = blk + wo + 8 * bq;
      int p = blk + wo;
      while (p < whi) {
        const int vhi = min(whi, run_end);  // points [p, vhi) belong to run `cur`
        half8 mhi, mlo;
        build_a<true>(hv, ev, pg0, p, vhi, mhi, mlo);
        mfma6<false>(mhi, mlo, cb[0], cb[1], cb[2], cb[3], acc);
        p = vhi;
        if (p == run_end) finish_run(p);
      }
      if constexpr (kGenNext) gen_slice(j, buf ^ 1);
      if constexpr (kNext) {
        load_eh(wo + kWin, ev, hv);
        load_b(wo + kWin, nb);
        build_a<false>(hv, ev, 0, 0, 0, na[0], na[1]);
      }
    }
  };
  auto first_window_at = [&](int wo, half8 (&bv)[4], half8 (&av)[2]) {
    float ev[8], hv[8];
    load_eh(wo, ev, hv);
    load_b(wo, bv);
    build_a<false>(hv, ev, 0, 0, 0, av[0], av[1]);
  };
  auto first_window = [&](half8 (&bv)[4], half8 (&av)[2]) { first_window_at(0, bv, av); };
  // a full block followed by another: kNWin windows, fully unrolled (the
  // operand register sets alternate), generating the next block
  auto windows_full = [&]() {
    half8 bx[4], by[4], ax[2], ay[2];
    first_window(bx, ax);
#pragma unroll
    for (int j = 0; j < kNWin; j += 2) {
      step(std::true_type{}, std::true_type{}, j * kWin, bx, ax, by, ay);
      if (j + 2 < kNWin)
        step(std::true_type{}, std::true_type{}, (j + 1) * kWin, by, ay, bx, ax);
      else
        step(std::true_type{}, std::false_type{}, (j + 1) * kWin, by, ay, bx, ax);
    }
  };
  // the last block of the chunk (possibly partial): nothing to generate
  auto windows_last = [&]() {
    half8 bx[4], ax[2];
#pragma unroll 1
    for (int wo = 0; wo < bn; wo += kWin) {
      first_window_at(wo, bx, ax);
      step(std::false_type{}, std::false_type{}, wo, bx, ax, bx, ax);
    }
  };

  // block 0 is generated up front; then block i+1 is generated while block i
  // is multiplied out.  A block that is followed by another in this chunk is
  // full (kBlk points, kNWin windows), so every slice of the next block's
  // generation gets its window.
  prefetch(b);
  gen_setup(0);
  write_h(0);
#pragma unroll 1
  for (int j = 0; j < kNWin; ++j) gen_slice(j, 0);
  if (b + kBlk < e) prefetch(b + kBlk);
  for (; blk < e; blk += kBlk, buf ^= 1) {
    // one barrier per block: block i is complete in `buf`, and every read of
    // the other buffer (block i-1) is done before block i+1 is written there
    __syncthreads();
    bn = min(kBlk, e - blk);
    if (blk + kBlk < e) {  // uniform
      gen_setup(buf ^ 1);
      if (blk + 2 * kBlk < e) prefetch(blk + 2 * kBlk);
      windows_full();
      write_h(buf ^ 1);
    } else {
      windows_last();
    }
  }
}

template <int M1, int KB, bool kGather>
__global__ __launch_bounds__(16 * M1, 2) void k_scatter_mfma(const PointRec* __restrict__ rec,
                                                          const int* __restrict__ seg_start,
                                                          int nseg, int chunk, int Nl,
                                                          const int* __restrict__ run_base,
                                                          const double* __restrict__ wsort,
                                                          float* __restrict__ part,
                                                          int* __restrict__ run_seg,
                                                          const unsigned* __restrict__ wmax) {
  const int c = blockIdx.x;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;
  scatter_mfma_item<M1, KB, kGather>(rec, seg_start, b, e, run_base[c], 0, Nl, blockIdx.y, wsort,
                                     part, run_seg, wmax);
}

// the work items of a group of point sets (dsht_transform_many): global
// chunk c belongs to set s = many_set_of(chunk_off, c)
template <int M1, int KB>
__global__ __launch_bounds__(16 * M1, 2) void k_scatter_mfma_many(
    ManySets ms, const PointRec* __restrict__ rec, const int* __restrict__ seg_start, int nseg,
    int stride, int Nl, const int* __restrict__ run_base, float* __restrict__ part,
    int* __restrict__ run_seg, const unsigned* __restrict__ wmax) {
  const int c = blockIdx.x;
  const int s = many_set_of(ms.chunk_off, ms.nsets, c);
  const int* ss = seg_start + s * stride;
  const int nvalid = ss[nseg];
  const int b = ms.off[s] + (c - ms.chunk_off[s]) * ms.chunk[s];
  const int e = min(nvalid, b + ms.chunk[s]);
  if (b >= e) return;
  scatter_mfma_item<M1, KB, false>(rec, ss, b, e, run_base[c], s * stride, Nl, blockIdx.y,
                                   nullptr, part, run_seg, wmax + s);
}

template <int M1, int KB>
static int launch_m1(Plan& pl, int nchunks, const double* w, bool gather, const unsigned* wmax,
                     hipStream_t s) {
  dim3 grid(nchunks, (pl.Nl + 32 * M1 - 1) / (32 * M1));
  if (gather)
    k_scatter_mfma<M1, KB, true><<<grid, 16 * M1, 0, s>>>(pl.ws.rec, pl.ws.seg_start, pl.nseg,
                                                         pl.chunk, pl.Nl, pl.ws.run_base, w,
                                                         (float*)pl.ws.part, pl.ws.run_seg, wmax);
  else
    k_scatter_mfma<M1, KB, false><<<grid, 16 * M1, 0, s>>>(pl.ws.rec, pl.ws.seg_start, pl.nseg,
                                                          pl.chunk, pl.Nl, pl.ws.run_base, w,
                                                          (float*)pl.ws.part, pl.ws.run_seg, wmax);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

template <int M1, int KB>
static int launch_many_m1(Plan& pl, const ManySets& ms, const unsigned* wmax, hipStream_t s) {
  const int nchunks = ms.chunk_off[ms.nsets];
  dim3 grid(nchunks, (pl.Nl + 32 * M1 - 1) / (32 * M1));
  k_scatter_mfma_many<M1, KB><<<grid, 16 * M1, 0, s>>>(ms, pl.ws.rec, pl.ws.seg_start, pl.nseg,
                                                      pl.nseg + 2, pl.Nl, pl.ws.run_base,
                                                      (float*)pl.ws.part, pl.ws.run_seg, wmax);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_scatter_mfma_many(Plan& pl, const ManySets& ms, const unsigned* wmax, hipStream_t s) {
  if (ms.chunk_off[ms.nsets] == 0) return DSHT_OK;
  if (pl.Nl <= 256) return launch_many_m1<8, 64>(pl, ms, wmax, s);
  if (pl.Nl <= 512) return launch_many_m1<16, 64>(pl, ms, wmax, s);
  return launch_many_m1<32, 128>(pl, ms, wmax, s);
}

// *wmax = max |w[i]| over the finite entries of one weight set (float bits).
// Four independent loads per thread and step.
__global__ __launch_bounds__(256) void k_absmax(int64_t n, const double* __restrict__ w,
                                                unsigned* __restrict__ wmax) {
  __shared__ float red[4];
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  float v[4] = {0.f, 0.f, 0.f, 0.f};
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = w[i + k * st];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float f = (float)fabs(a[k]);
      if (f <= 3.0e38f) v[k] = fmaxf(v[k], f);
    }
  }
  for (; i < n; i += st) {
    const float f = (float)fabs(w[i]);
    if (f <= 3.0e38f) v[0] = fmaxf(v[0], f);
  }
  float m = fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    atomic_max_pos(wmax, b);
  }
}

// p[0 .. n) = 0 and, if q, q[0 .. n) = 0
__global__ void k_clear_u32(unsigned* __restrict__ p, unsigned* __restrict__ q, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    p[i] = 0u;
    if (q) q[i] = 0u;
  }
}

int launch_clear_u32(unsigned* p, unsigned* q, int n, hipStream_t s) {
  if (n <= 0) return DSHT_OK;
  k_clear_u32<<<(n + 255) / 256, 256, 0, s>>>(p, q, n);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

// m tile of 256, 512 or 1024 orders: the narrowest that covers Nl, or 1024.
// The operand scale comes from ws.wmax[0]: the points pass's max |w| for the
// records' own weights; for a gathered weight set, its own max, reduced here.
int launch_weight_max(Plan& pl, int64_t Np, const double* w, hipStream_t s, unsigned* out) {
  if (!out) out = pl.ws.wmax;
  int rc;
  if ((rc = launch_clear_u32(out, nullptr, 1, s))) return rc;
  if (Np <= 0) return DSHT_OK;
  // up to 2048 blocks, four loads in flight per thread
  const int blocks = (int)std::min<int64_t>((Np + 1023) / 1024, 2048);
  k_absmax<<<blocks, 256, 0, s>>>(Np, w, out);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_scatter_mfma(Plan& pl, int64_t Np, const double* w, bool gather, hipStream_t s) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  const unsigned* wmax = pl.ws.wmax;
  int rc;
  // records hold the basis alone (unit weights): the set's weights set the range
  if ((gather || w) && (rc = launch_weight_max(pl, Np, w, s))) return rc;
  if (pl.Nl <= 256) return launch_m1<8, 64>(pl, nchunks, w, gather, wmax, s);
  if (pl.Nl <= 512) return launch_m1<16, 64>(pl, nchunks, w, gather, wmax, s);
  return launch_m1<32, 128>(pl, nchunks, w, gather, wmax, s);
}

}  // namespace dsht
