// Stage (b), part 2, for many weight sets on shared points (DSHT_FP32 plans,
// dsht_transform_sets): the scatter of all sets at once,
//   v^s_k(m, run) = sum_{p in run} w_s(p) h_k(p) e^{-i m phi_p},
// the v's of sht/shared_interp_funcs.py:88-105 once per weight set, as one
// matrix-core GEMM per run with the sets as rows:
//   C[s][(k, m, re|im)] = sum_p W[s][p] B[p][(k, m, re|im)],
//   W[s][p] = w_s(p),  B[p][(k, m, c)] = h_k(p) (Re | Im) e^{-i m phi_p}.
// Each B entry is formed once for all sets (and each weight split once for
// all orders), so the vector work per matrix-core product is ~4 instructions
// (against ~10 for the per-set scatter), and the real-times-complex product
// is 2 flop per (set, point, m, k) (the per-set scatter's real GEMM form of a
// complex product takes 4).  Both operands are split into fp16 hi + lo and
// three products hi*hi + hi*lo + lo*hi are accumulated in FP32 by
// v_mfma_f32_32x32x16_f16, as in the per-set scatter.
//
// e^{-i m phi} = E(m1) B(m2), m = 32 m1 + m2, with E and B from FP64 chains
// (steps e1 = e^{-i phi} and e1^32) rounded to FP32; their FP32 complex
// product is within ~1e-7 of e^{-i m phi}.
//
// Each weight set is scaled by a power of two 2^-e_s that brings max |w_s|
// into [0.5, 1) (exact; undone at the run's write), so the fp16 split keeps
// ~2^-22 relative precision whatever the weights' magnitude.
//
// The weights are first put in sorted order once (FP32 rows wsf[s][pos], by a
// transpose to point-major rows and a gather of whole rows, which also finds
// max |w_s|), so the scatter reads them coalesced.
//
// Two kernels: k_scatter_sets below (sets as rows, 33..64 sets) and
// k_scatter_wsets further down (rows (k, set), up to 32 sets), chosen in
// launch_scatter_sets by the set count on measured times.
//
// k_scatter_sets workgroups: (chunk of sorted points, block of 64 orders), the order blocks
// of a chunk on one XCD.  One workgroup of 4 wavefronts, two per CU (each
// wavefront's 128 accumulator registers use half of a SIMD's register file);
// wavefront w owns 16 orders m = 64 my + 16 w + (0..15), i.e. four 32-column
// tiles (one per k, columns (m, re|im)), and RT row tiles of 32 sets.  The
// chunk is staged kBlk points at a time: the sets' weights (scaled, split
// into fp16 hi/lo), E for the workgroup's two m1, B for all 32 m2, and h_k;
// one workgroup's staging overlaps the other's products.  A 16-point window
// (one MFMA K step) is issued once per run it meets (once for almost all
// windows), with the points outside the run zeroed.  A run's result goes to
// its own slot part[set][run][k][m][2] (FP32), the layout the fold reads for
// each set.
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

constexpr int kBlk = 64;         // points per staged block (a chunk is a multiple of it)
constexpr int kRow = kBlk + 8;   // fp16 rows: 4 (mod 32) dwords
constexpr int kRowF = kBlk + 4;  // FP32 rows: 4 (mod 32) dwords
constexpr int kThreads = 256;    // 4 wavefronts: 64 orders
constexpr int kOrders = 64;      // orders per workgroup
constexpr int kWin = 16;         // points per MFMA K step

}  // namespace

__global__ void k_fill_u32(unsigned* __restrict__ p, int n, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// sidx[pos] = input index of sorted point pos (valid points only)
__global__ __launch_bounds__(256) void k_sorted_idx(int64_t Np, const int* __restrict__ nvalid,
                                                    const PointRec* __restrict__ rec,
                                                    int* __restrict__ sidx) {
  const int64_t pos = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (pos < Np && pos < *nvalid) sidx[pos] = rec[pos].idx;
}

// The sort of the weight sets into the scatter's order is two coalesced
// passes instead of nsets random 8-byte gathers:
//  k_transpose_w: wt[idx][s] = (float) w_s[idx]  (point-major, nsets floats a
//                 point), 64 x 64 tiles through LDS;
//  k_gather_sets: wsf[s][pos] = wt[sidx[pos]][s] (one 4 * nsets-byte row per
//                 sorted point), zero past the valid points up to n, again
//                 through a 64 x 64 LDS tile, and smax[s] = max |w_s| as float
//                 bits (non-negative floats order like their bit patterns;
//                 smax cleared first).  A block covers kGatherTiles tiles and
//                 issues one atomic per set (same-address atomics serialize).
constexpr int kTile = 64;
constexpr int kGatherTiles = 16;
__global__ __launch_bounds__(256) void k_transpose_w(const double* __restrict__ w, int64_t ldw,
                                                     int64_t Np, int nsets,
                                                     float* __restrict__ wt) {
  __shared__ float tile[kTile][kTile + 1];  // [idx][set]
  const int64_t i0 = (int64_t)blockIdx.x * kTile;
  const int s0 = blockIdx.y * kTile;
  const int t = threadIdx.x, lo = t & 63, hi = t >> 6;
#pragma unroll
  for (int r = 0; r < kTile / 4; ++r) {
    const int sl = hi + 4 * r;
    const int64_t i = i0 + lo;
    tile[lo][sl] = (s0 + sl < nsets && i < Np) ? (float)w[(int64_t)(s0 + sl) * ldw + i] : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kTile / 4; ++r) {
    const int il = hi + 4 * r;
    const int64_t i = i0 + il;
    if (s0 + lo < nsets && i < Np) wt[i * nsets + s0 + lo] = tile[il][lo];
  }
}

__global__ __launch_bounds__(256) void k_gather_sets(const float* __restrict__ wt, int nsets,
                                                     const int* __restrict__ nvalid,
                                                     const int* __restrict__ sidx, int64_t n,
                                                     float* __restrict__ wsf, int64_t ldws,
                                                     unsigned* __restrict__ smax) {
  __shared__ float tile[kTile][kTile + 1];  // [pos][set]
  const int nv = *nvalid;
  const int t = threadIdx.x, lo = t & 63, hi = t >> 6;
  float mx[kTile / 4];
#pragma unroll
  for (int r = 0; r < kTile / 4; ++r) mx[r] = 0.f;
  for (int tt = 0; tt < kGatherTiles; ++tt) {
    const int64_t p0 = ((int64_t)blockIdx.x * kGatherTiles + tt) * kTile;
    if (p0 >= n) break;  // uniform
#pragma unroll
    for (int r = 0; r < kTile / 4; ++r) {
      const int pl = hi + 4 * r;
      const int64_t pos = p0 + pl;
      float v = 0.f;
      if (pos < nv && lo < nsets) v = wt[(int64_t)sidx[pos] * nsets + lo];
      tile[pl][lo] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kTile / 4; ++r) {
      const int sl = hi + 4 * r;
      const int64_t pos = p0 + lo;
      const float v = tile[lo][sl];
      if (sl < nsets && pos < n) wsf[(int64_t)sl * ldws + pos] = v;
      mx[r] = fmaxf(mx[r], fabsf(v));
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < kTile / 4; ++r) {
    float m = mx[r];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    const int sl = hi + 4 * r;
    if (lo == 0 && sl < nsets) atomic_max_pos(smax + sl, m);
  }
}

// The node spreading's weight sets, point-major: k_place_w writes
// wt[i nsets + s] = (float) w_s[i], one 64 x 64 tile (points x sets) through
// LDS per step (coalesced reads of each set's weights, each point's row of
// nsets floats written whole), and wmax[s] = max |w_s| over the finite
// weights (of every input point: an upper bound for the points the binning
// keeps), one atomic per set per block.  The node kernel reads a point's
// weights through its record's input index: the workgroups of all sets of a
// node pair run together on one XCD, so each row comes from HBM once.
__global__ __launch_bounds__(256) void k_place_w(const double* __restrict__ w, int64_t ldw,
                                                 int64_t Np, int nsets,
                                                 float* __restrict__ wt,
                                                 unsigned* __restrict__ wmax) {
  __shared__ float tile[kTile][kTile + 1];  // [point][set]
  const int t = threadIdx.x, lo = t & 63, hi = t >> 6;
  float mx[kTile / 4];
#pragma unroll
  for (int r = 0; r < kTile / 4; ++r) mx[r] = 0.f;
  for (int tt = 0; tt < kGatherTiles; ++tt) {
    const int64_t i0 = ((int64_t)blockIdx.x * kGatherTiles + tt) * kTile;
    if (i0 >= Np) break;  // uniform
#pragma unroll
    for (int r = 0; r < kTile / 4; ++r) {
      const int sl = hi + 4 * r;
      const int64_t i = i0 + lo;
      float v = 0.f;
      if (sl < nsets && i < Np) v = (float)w[(int64_t)sl * ldw + i];
      tile[lo][sl] = v;
      const float a = fabsf(v);
      if (a <= 3.0e38f) mx[r] = fmaxf(mx[r], a);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kTile / 4; ++r) {
      const int il = hi + 4 * r;
      const int64_t i = i0 + il;
      if (i < Np && lo < nsets) wt[i * nsets + lo] = tile[il][lo];
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < kTile / 4; ++r) {
    float m = mx[r];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    const int sl = hi + 4 * r;
    if (lo == 0 && sl < nsets) atomic_max_pos(wmax + sl, m);
  }
}

int launch_place_weight_sets(Plan& pl, int64_t Np, const double* w, int64_t ldw, int nsets,
                            hipStream_t s) {
  Workspace& ws = pl.ws;
  if (nsets < 1 || nsets > 64 || nsets > ws.cap_sets || nsets > ws.cap_wsf_sets || !ws.wsf ||
      Np > ws.cap_set_points) {
    set_error("dsht: placing 1..64 weight sets within the workspace");
    return DSHT_EINVAL;
  }
  k_fill_u32<<<1, 64, 0, s>>>(ws.wmax, nsets, 0u);
  DSHT_LAUNCH_CHECK();
  if (Np == 0) return DSHT_OK;
  const int64_t per = (int64_t)kTile * kGatherTiles;
  k_place_w<<<(unsigned)((Np + per - 1) / per), 256, 0, s>>>(w, ldw, Np, nsets, ws.wsf, ws.wmax);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

// Weight-set scatter, RT row tiles of 32 sets (RT = 2: up to 64 sets).
template <int RT>
__global__ __launch_bounds__(kThreads, 1) void k_scatter_sets(
    const PointRec* __restrict__ rec, const int* __restrict__ seg_start, int nseg, int chunk, int Nl,
    const int* __restrict__ run_base, const float* __restrict__ wsf, int64_t ldws, int nsets,
    const unsigned* __restrict__ smax, int nchunks, int nmy, float* __restrict__ part,
    int64_t set_stride, int* __restrict__ run_seg) {
  // workgroup b runs on XCD b % 8; the nmy order blocks of one chunk go to
  // one XCD, one after another, so they share its L2 for the chunk's records
  // and weights
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int c = 8 * (slot / nmy) + xcd, my = slot % nmy;
  if (c >= nchunks) return;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;

  // double-buffered: a block's weights (FP32, unscaled; filled by LDS DMA,
  // global_load_lds) and its z, h and segments (computed by the threads);
  // block i+1's are written while block i is multiplied out
  __shared__ __attribute__((aligned(16))) float sWr[2][32 * RT][kRowF];   // [buf][set][point]
  __shared__ __attribute__((aligned(16))) float sZ[2][2 * kOrders][kRowF];  // [buf][(m, re|im)][point]
  __shared__ __attribute__((aligned(16))) float sH[2][4][kBlk];
  __shared__ int sSeg[2][kBlk];
  __shared__ float sInv[32 * RT];  // 2^e_s

  // weight set s is scaled by 2^-e_s, max |w_s| 2^-e_s in [0.5, 1); this
  // lane's A rows are sets 32 rt + l % 32
  float asc[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    const int s = 32 * rt + (l & 31);
    int ex = 0;
    if (s < nsets) {
      const float mx = __uint_as_float(smax[s]);
      if (mx > 0.f && mx < 3.0e38f) (void)frexpf(mx, &ex);
    }
    asc[rt] = ldexpf(1.0f, -ex);
    if (wv == 0 && l < 32) sInv[s] = ldexpf(1.0f, ex);
  }

  int blk = b, bn = 0, buf = 0;  // current block, its number of points, its buffer

  // LDS DMA of the weights of the block at pblk into buffer nb: one 256-byte
  // row per instruction, sets wv + 4 t below nsets (rows of other sets are
  // never used: their results are not written).  wsf rows are long enough
  // for a whole block.
  auto dma_w = [&](int pblk, int nb) {
#pragma unroll
    for (int t = 0; t < 8 * RT; ++t) {
      const int s = wv + 4 * t;
      if (s < nsets) {
        const float* row = wsf + s * ldws + pblk;  // uniform
        __builtin_amdgcn_global_load_lds(row + l, &sWr[nb][s][0], 4, 0, 0);
      }
    }
  };
  // staging: thread (point gp, group gg) computes z(m, p) = e^{-i m phi_p}
  // for orders 16 gg .. 16 gg + 15 of the workgroup by an FP64 chain from
  // e1^(64 my + 16 gg), and h_gg; the chain runs in four slices of four
  // orders
  const int gp = tid & (kBlk - 1);
  const int gg = __builtin_amdgcn_readfirstlane(tid / kBlk);
  struct RecF {
    double re, im, h;
    int seg;
  };
  // this thread's record fields of the block at pblk (padding points repeat
  // the block's first record)
  auto load_rec = [&](int pblk) {
    const int pbn = min(kBlk, e - pblk);
    const PointRec* pr = rec + pblk + (gp < pbn ? gp : 0);
    RecF f;
    f.re = pr->ere;
    f.im = pr->eim;
    f.h = gg == 0 ? pr->h0 : gg == 1 ? pr->h1 : gg == 2 ? pr->h2 : pr->h3;
    f.seg = pr->seg;
    return f;
  };
  auto zbase = [&](cplx e1) {
    cplx z{1.0, 0.0}, pw = e1;
    int nn = kOrders * my + 16 * gg;
    while (nn) {
      if (nn & 1) z = cmul(z, pw);
      pw = csq(pw);
      nn >>= 1;
    }
    return z;
  };
  // orders 16 gg + q0 .. + 3 from the chain state zs (= z at q0) into buffer
  // nb; returns the state at q0 + 4
  auto zslice = [&](int nb, cplx e1, cplx zs, int q0) {
    cplx z = zs;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sZ[nb][2 * (16 * gg + q0 + q)][gp] = (float)z.x;
      sZ[nb][2 * (16 * gg + q0 + q) + 1][gp] = (float)z.y;
      z = cmul(z, e1);
    }
    return z;
  };
  // h and segment of the block in buffer nb (the segment written by all four
  // groups, with the same value)
  auto stage_hs = [&](int nb, const RecF& f, bool ok) {
    sH[nb][gg][gp] = ok ? (float)f.h : 0.f;
    sSeg[nb][gp] = ok ? f.seg : -1;
  };

  // this lane's column: m = 64 my + 16 wv + (col >> 1), component col & 1
  const int col = l & 31, hq = l >> 5;
  const int m = kOrders * my + 16 * wv + (col >> 1);
  const int cc = col & 1;
  const int zrow = 32 * wv + col;

  floatx16 acc[RT][4];
  const floatx16 zero = {};
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[rt][k] = zero;

  int run = __builtin_amdgcn_readfirstlane(run_base[c]);
  int cur = __builtin_amdgcn_readfirstlane(rec[b].seg);
  int run_end = min(e, seg_start[cur + 1]);

  // the current run ends at point p: write its slot (unscaled), clear the
  // accumulators, move to the next run.  Lane l, register r of tile rt holds
  // C[set 32 rt + 8 (r/4) + 4 (l/32) + r%4][column l%32].
  auto finish_run = [&](int p) {
    if (m < Nl) {
      // the accumulators are read with volatile inline asm, which stays in
      // this branch (the compiler would otherwise hoist the AGPR reads into
      // every pass).  The compiler does not see these reads as dependent on
      // the MFMAs that wrote the registers, so first wait out the longest
      // MFMA result latency (3 x 8 wait states).
      asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" ::: "memory");
      float* q = part + ((int64_t)run * 4 * Nl + m) * 2 + cc;
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int s = 32 * rt + 8 * (r >> 2) + 4 * hq + (r & 3);
          if (s < nsets) {
            const float inv = sInv[s];
            float* qs = q + s * set_stride;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              float v;
              asm volatile("v_accvgpr_read_b32 %0, %1" : "=v"(v) : "a"(acc[rt][k][r]));
              qs[(int64_t)k * Nl * 2] = v * inv;
            }
          }
        }
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[rt][k] = zero;
    if (tid == 0 && my == 0) run_seg[run] = cur;
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

  // block b is staged up front
  dma_w(b, 0);
  {
    const RecF f = load_rec(b);
    const bool ok = gp < min(kBlk, e - b);
    const cplx e1 = ok ? cplx{f.re, f.im} : cplx{1.0, 0.0};
    cplx z = zbase(e1);
#pragma unroll
    for (int q0 = 0; q0 < 16; q0 += 4) z = zslice(0, e1, z, q0);
    stage_hs(0, f, ok);
  }
  RecF fn{1.0, 0.0, 0.0, -1};  // record fields of the next block
  if (b + kBlk < e) fn = load_rec(b + kBlk);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();

  for (; blk < e; blk += kBlk, buf ^= 1) {
    bn = min(kBlk, e - blk);
    // the next block: its weights by DMA now, its z, h and segments in
    // slices interleaved with this block's products (computed for a
    // nonexistent next block too, into the idle buffer, which keeps the
    // product code free of branches); the block after it: record fields
    const int nblk = blk + kBlk;
    const int nbn = min(kBlk, e - nblk);  // <= 0 if there is no next block
    if (nblk < e) dma_w(nblk, buf ^ 1);
    RecF fnn{1.0, 0.0, 0.0, -1};
    if (nblk + kBlk < e) fnn = load_rec(nblk + kBlk);
    const bool nok = gp < nbn;
    const cplx ne1 = nok ? cplx{fn.re, fn.im} : cplx{1.0, 0.0};
    cplx zst = ne1;  // chain state between slices
#pragma unroll
    for (int w = 0; w < kBlk / kWin; ++w) {
      const int wo = kWin * w;
      if (wo < bn) {  // uniform
        const int wlo = blk + wo, whi = blk + min(wo + kWin, bn);
        const int po = wo + 8 * hq;  // this lane's K slice: points po .. po+7
        float zv[8];
        *reinterpret_cast<float4*>(&zv[0]) = *reinterpret_cast<const float4*>(&sZ[buf][zrow][po]);
        *reinterpret_cast<float4*>(&zv[4]) =
            *reinterpret_cast<const float4*>(&sZ[buf][zrow][po + 4]);
        half8 whi_[RT], wlo_[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          float wa[8];
          const float* wr = &sWr[buf][32 * rt + (l & 31)][po];
          *reinterpret_cast<float4*>(&wa[0]) = *reinterpret_cast<const float4*>(wr);
          *reinterpret_cast<float4*>(&wa[4]) = *reinterpret_cast<const float4*>(wr + 4);
#pragma unroll
          for (int j = 0; j < 8; ++j) wa[j] *= asc[rt];
          split8_f16(wa, whi_[rt], wlo_[rt]);
        }
        // one pass per run the window meets (one pass for almost all
        // windows): this window's 4 * 3 * RT products over the points
        // [p, vhi) of run `cur`, the other points' z zeroed, and slice w of
        // the next block's staging (repeated identically by a second pass).
        // A single code path for all windows keeps each accumulator in one
        // register set.
        cplx zout = zst;
        int p = wlo;
        while (p < whi) {
          const int vhi = min(whi, run_end);
          float zm[8];
          const int rel0 = blk + po - p;
#pragma unroll
          for (int j = 0; j < 8; ++j)
            zm[j] = (unsigned)(rel0 + j) < (unsigned)(vhi - p) ? zv[j] : 0.f;
          half8 bh[4], bl[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = sH[buf][k][po + j] * zm[j];
            split8_f16(x, bh[k], bl[k]);
          }
          if (w == 0) {  // compile-time
            zout = zslice(buf ^ 1, ne1, zbase(ne1), 0);
            stage_hs(buf ^ 1, fn, nok);
          } else {
            zout = zslice(buf ^ 1, ne1, zst, 4 * w);
          }
          // 3 rounds over the 4 RT independent accumulators, so that an
          // MFMA never waits for the one just issued
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
              acc[rt][k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(whi_[rt], bh[k], acc[rt][k], 0, 0, 0);
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
              acc[rt][k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(whi_[rt], bl[k], acc[rt][k], 0, 0, 0);
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
              acc[rt][k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(wlo_[rt], bh[k], acc[rt][k], 0, 0, 0);
          p = vhi;
          if (p == run_end) finish_run(p);
          // keep the accumulators in AGPRs across the loop (otherwise the
          // compiler carries them in VGPRs and copies all of them every pass)
#pragma unroll
          for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int k = 0; k < 4; ++k) asm volatile("" : "+a"(acc[rt][k]));
        }
        zst = zout;
      }
    }
    fn = fnn;
    // the next block's weights have landed (this wavefront's DMA) and, after
    // the barrier, its staging is complete and this block's readers are done
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Weight-set scatter with the Hermite index on the set side:
//   C[(k, s)][(m, re|im)] = sum_p A[(k, s)][p] Z[p][(m, re|im)],
//   A[(k, s)][p] = w_s(p) h_k(p) 2^-e_s,   Z[p][(m, c)] = (Re | Im) e^{-i m phi_p}.
// A has 4 SP rows (SP = 8 RTA sets, padded), built and split into fp16 hi/lo
// once per staged block in LDS and read by every wavefront; Z is built once
// per block in LDS in FP32 and each wavefront splits only its own columns.
// So the vector work per matrix-core product is ~3-4 instructions (the A
// products and splits are shared by 128 orders, the Z chains by 4 SP rows),
// against ~7 when each wavefront forms h_k z for its orders and splits every
// weight itself.
//
// Workgroup: (chunk of sorted points, block of 128 orders), NW wavefronts;
// wavefront w owns orders 128 my + 16 CT w + (0 .. 16 CT - 1) as CT 32-column
// tiles, and all RTA row tiles: RTA CT accumulator tiles of 16 registers,
// independent, so consecutive MFMAs never wait on each other.  Blocks of
// kWBlk = 32 points: while block i is multiplied out, block i+1's A and Z are
// built into the other buffer from registers loaded at the start of the
// block; one barrier per block.
// ---------------------------------------------------------------------------
constexpr int kWBlk = 32;                 // points per staged block
constexpr int kWPg = kWBlk / 8;           // 8-point groups (MFMA K slices) per block
constexpr int kWOrders = 128;             // orders per workgroup
// Z in LDS, FP32: value (point p, column col) at
//   (2 pg + hf) * kZPg + (col + col / 32) * 4 + c4,   p = 8 pg + 4 hf + c4,
// so a lane's 8 K values are two 16-byte reads, the reads of 32 consecutive
// columns are contiguous, and the stores of 32 points at one column hit 32
// different banks (kZPg = 4 mod 32).
constexpr int kZPg = 1060;
constexpr int kZBuf = 8 * kZPg;           // floats per buffer

__device__ inline int zidx(int p, int col) {
  return ((p >> 2) * kZPg) + (col + (col >> 5)) * 4 + (p & 3);
}

template <int RTA, int CT, int NW>
__global__ __launch_bounds__(64 * NW, 1) void k_scatter_wsets(
    const PointRec* __restrict__ rec, const int* __restrict__ seg_start, int nseg, int chunk,
    int Nl, const int* __restrict__ run_base, const float* __restrict__ wsf, int64_t ldws,
    int nsets, const unsigned* __restrict__ smax, int nchunks, int nmy, float* __restrict__ part,
    int64_t set_stride, int* __restrict__ run_seg) {
  static_assert(16 * CT * NW == kWOrders, "the wavefronts cover 128 orders");
  constexpr int kT = 64 * NW;           // threads
  constexpr int SP = 8 * RTA;           // sets per row block (padded)
  constexpr int NR = 4 * SP;            // A rows, r = k SP + s
  constexpr int kElems = NR * kWPg;     // A elements (row, 8-point group) per block
  constexpr int kPerThr = (kElems + kT - 1) / kT;
  constexpr int kZG = kT / kWBlk;       // z chain groups per point
  constexpr int kZO = kWOrders / kZG;   // orders per chain
  constexpr int kRP = RTA >= 2 ? 2 : 1;  // row tiles per MFMA round
  // workgroup b runs on XCD b % 8; the order blocks of one chunk go to one
  // XCD, one after another, so they share its L2 for the chunk's records
  // and weights
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int c = 8 * (slot / nmy) + xcd, my = slot % nmy;
  if (c >= nchunks) return;
  const int tid = threadIdx.x, l = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;

  __shared__ __attribute__((aligned(16))) half8 sA[2][2][kWPg][NR];  // [buf][hi|lo][pg][row]
  __shared__ __attribute__((aligned(16))) float sZ[2][kZBuf];
  __shared__ __attribute__((aligned(16))) float sH[2][4][kWBlk];     // h_k of a block's points
  __shared__ float sScale[2][SP];                                    // 2^-e_s, 2^e_s

  if (tid < SP) {
    int ex = 0;
    if (tid < nsets) {
      const float mx = __uint_as_float(smax[tid]);
      if (mx > 0.f && mx < 3.0e38f) (void)frexpf(mx, &ex);
    }
    sScale[0][tid] = ldexpf(1.0f, -ex);
    sScale[1][tid] = ldexpf(1.0f, ex);
  }

  // ---- staging roles
  // z chains: thread (point zp, group zg) makes orders kZO zg .. kZO zg + kZO - 1;
  // h loads: thread (point zp, k = zg) for zg < 4;
  // A elements: element q = tid + kT j (j < kPerThr): row q % NR, group q / NR
  const int zp = tid & (kWBlk - 1), zg = tid / kWBlk;
  double se_re = 1.0, se_im = 0.0;  // e^{-i phi} of point zp of the block being staged
  float4 sw[kPerThr][2];            // weights of element j's 8 points (unscaled)
  double pe_re = 1.0, pe_im = 0.0;  // the same, loaded a block earlier (prefetch)
  float4 pw[kPerThr][2];
  float ph = 0.f;                   // h_zg of point zp of a later block
  // records and weights of the block at pblk into the prefetch registers
  // (past e: zeros); take() moves them to the staging registers
  auto load_stage = [&](int pblk) {
    const int p = pblk + zp;
    const bool ok = p < e;
    const PointRec* pr = rec + (ok ? p : b);
    pe_re = ok ? pr->ere : 1.0;
    pe_im = ok ? pr->eim : 0.0;
#pragma unroll
    for (int j = 0; j < kPerThr; ++j) {
      const int q = tid + kT * j;
      const int row = q % NR, pg = q / NR;
      const int s = row % SP;
      float4 a = {0.f, 0.f, 0.f, 0.f}, bq = {0.f, 0.f, 0.f, 0.f};
      if (q < kElems && s < nsets && pblk < e) {
        // rows hold every position a block reads, zero past the valid points
        const float4* wr = reinterpret_cast<const float4*>(wsf + s * ldws + pblk + 8 * pg);
        a = wr[0];
        bq = wr[1];
      }
      pw[j][0] = a;
      pw[j][1] = bq;
    }
  };
  auto take = [&]() {
    se_re = pe_re;
    se_im = pe_im;
#pragma unroll
    for (int j = 0; j < kPerThr; ++j) {
      sw[j][0] = pw[j][0];
      sw[j][1] = pw[j][1];
    }
  };
  auto load_h = [&](int pblk) {
    const int p = pblk + zp;
    if (zg < 4) {
      float v = 0.f;
      if (p < e) {
        const PointRec* pr = rec + p;
        v = (float)(zg == 0 ? pr->h0 : zg == 1 ? pr->h1 : zg == 2 ? pr->h2 : pr->h3);
      }
      ph = v;
    }
  };
  // A elements j0 .. j1-1 of the staged block into buffer nb (h from sH[nb])
  auto stage_a = [&](int nb, int j0, int j1) {
#pragma unroll
    for (int j = j0; j < j1; ++j) {
      const int q = tid + kT * j;
      if (q < kElems) {
        const int row = q % NR, pg = q / NR;
        const int k = row / SP, s = row % SP;
        const float sc = sScale[0][s];
        const float4 h0 = *reinterpret_cast<const float4*>(&sH[nb][k][8 * pg]);
        const float4 h1 = *reinterpret_cast<const float4*>(&sH[nb][k][8 * pg + 4]);
        float x[8] = {sw[j][0].x * h0.x, sw[j][0].y * h0.y, sw[j][0].z * h0.z, sw[j][0].w * h0.w,
                      sw[j][1].x * h1.x, sw[j][1].y * h1.y, sw[j][1].z * h1.z, sw[j][1].w * h1.w};
        uint32_t hi[4], lo[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) split2_f16(x[2 * u] * sc, x[2 * u + 1] * sc, hi[u], lo[u]);
        sA[nb][0][pg][row] = __builtin_bit_cast(half8, (u32x4){hi[0], hi[1], hi[2], hi[3]});
        sA[nb][1][pg][row] = __builtin_bit_cast(half8, (u32x4){lo[0], lo[1], lo[2], lo[3]});
      }
    }
  };
  // z chain of the staged block: seed at order 128 my + kZO zg, then steps
  cplx zst{1.0, 0.0};
  auto z_seed = [&]() {
    const cplx e1{se_re, se_im};
    cplx z{1.0, 0.0}, pw2 = e1;
    int nn = kWOrders * my + kZO * zg;
    while (nn) {
      if (nn & 1) z = cmul(z, pw2);
      pw2 = csq(pw2);
      nn >>= 1;
    }
    zst = z;
  };
  auto z_steps = [&](int nb, int q0, int q1) {
    const cplx e1{se_re, se_im};
    float* zb = sZ[nb];
#pragma unroll
    for (int q = q0; q < q1; ++q) {
      const int zc = 2 * (kZO * zg + q);
      zb[zidx(zp, zc)] = (float)zst.x;
      zb[zidx(zp, zc + 1)] = (float)zst.y;
      zst = cmul(zst, e1);
    }
  };

  // ---- products
  const int col = l & 31, hq = l >> 5;
  floatx16 acc[RTA][CT];
  const floatx16 zero = {};
#pragma unroll
  for (int rt = 0; rt < RTA; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = zero;

  int run = __builtin_amdgcn_readfirstlane(run_base[c]);
  int cur = __builtin_amdgcn_readfirstlane(rec[b].seg);
  int run_end = min(e, seg_start[cur + 1]);

  // lane l, register r of tile (rt, ct) holds C[row 32 rt + 8 (r/4) + 4 (l/32) + r%4]
  // [column 32 ct + l%32]: order 128 my + 16 (CT wv + ct) + (l%32)/2, component l%2
  // Row 32 rt + 8 (r/4) + 4 (l/32) + r%4 = k SP + s: per lane, s = s_lane +
  // (32 rt) % SP + 8 (r/4) + r%4 with s_lane = 4 (l/32), and k = (32 rt +
  // 8 (r/4)) / SP, so every store is this lane's base plus a uniform offset.
  auto finish_run = [&](int p) {
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" ::: "memory");
    int s_lane = 4 * hq;
    asm volatile("" : "+v"(s_lane));  // keeps the per-lane base inside this branch
    const float* scl = &sScale[1][s_lane];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int m = kWOrders * my + 16 * (CT * wv + ct) + (col >> 1);
      if (m < Nl) {
        float* q = part + ((int64_t)run * 4 * Nl + m) * 2 + (col & 1) + s_lane * set_stride;
#pragma unroll
        for (int rt = 0; rt < RTA; ++rt)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int row0 = 32 * rt + 8 * (r >> 2) + (r & 3);  // row less 4 (l/32)
            const int k = row0 / SP, sc = row0 % SP;
            if (sc + s_lane < nsets) {
              float v;
              asm volatile("v_accvgpr_read_b32 %0, %1" : "=v"(v) : "a"(acc[rt][ct][r]));
              q[sc * set_stride + (int64_t)k * Nl * 2] = v * scl[sc];
            }
          }
      }
    }
#pragma unroll
    for (int rt = 0; rt < RTA; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = zero;
    if (tid == 0 && my == 0) run_seg[run] = cur;
    if (p < e) {
      ++run;
      cur = __builtin_amdgcn_readfirstlane(rec[p].seg);
      run_end = min(e, seg_start[cur + 1]);
    }
  };
  auto pin_acc = [&]() {
#pragma unroll
    for (int rt = 0; rt < RTA; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) asm volatile("" : "+a"(acc[rt][ct]));
  };
  // one pass of window wdx of buffer nb over points [plo, phi_) of the
  // current run (the whole window when not masked): Z split once, then the
  // row tiles in rounds of kRP with their A slices loaded just before
  auto mfma_pass = [&](int nb, int wdx, int blk, bool masked, int plo, int phi_) {
    const int pg = 2 * wdx + hq;
    half8 bh[CT], bl[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int zc = 32 * (CT * wv + ct) + col;
      const float* zr = &sZ[nb][(2 * pg) * kZPg + (zc + (zc >> 5)) * 4];
      const float4 z0 = *reinterpret_cast<const float4*>(zr);
      const float4 z1 = *reinterpret_cast<const float4*>(zr + kZPg);
      float x[8] = {z0.x, z0.y, z0.z, z0.w, z1.x, z1.y, z1.z, z1.w};
      if (masked) {
        const int rel = blk + 8 * pg - plo;
        const unsigned len = (unsigned)(phi_ - plo);
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = (unsigned)(rel + u) < len ? x[u] : 0.f;
      }
      split8_f16(x, bh[ct], bl[ct]);
    }
#pragma unroll
    for (int r0 = 0; r0 < RTA; r0 += kRP) {
      half8 ah[kRP], al[kRP];
#pragma unroll
      for (int u = 0; u < kRP; ++u) {
        ah[u] = sA[nb][0][pg][32 * (r0 + u) + col];
        al[u] = sA[nb][1][pg][32 * (r0 + u) + col];
      }
#pragma unroll
      for (int u = 0; u < kRP; ++u)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
          acc[r0 + u][ct] =
              __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[u], bh[ct], acc[r0 + u][ct], 0, 0, 0);
#pragma unroll
      for (int u = 0; u < kRP; ++u)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
          acc[r0 + u][ct] =
              __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[u], bl[ct], acc[r0 + u][ct], 0, 0, 0);
#pragma unroll
      for (int u = 0; u < kRP; ++u)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
          acc[r0 + u][ct] =
              __builtin_amdgcn_mfma_f32_32x32x16_f16(al[u], bh[ct], acc[r0 + u][ct], 0, 0, 0);
    }
  };

  // ---- prologue: block 0 staged into buffer 0; block 1's h in sH[1] and
  // its records and weights in the staging registers; block 2's loading
  load_stage(b);
  load_h(b);
  if (zg < 4) sH[0][zg][zp] = ph;
  take();
  load_stage(b + kWBlk);
  __syncthreads();  // sH[0], sScale
  stage_a(0, 0, kPerThr);
  z_seed();
  z_steps(0, 0, kZO);
  load_h(b + kWBlk);
  if (zg < 4) sH[1][zg][zp] = ph;
  take();
  load_stage(b + 2 * kWBlk);
  __syncthreads();

  int buf = 0;
  for (int blk = b; blk < e; blk += kWBlk, buf ^= 1) {
    const bool next = blk + kWBlk < e;  // uniform: stage the next block
    const int bn = min(kWBlk, e - blk);
    load_h(blk + 2 * kWBlk);
#pragma unroll
    for (int wdx = 0; wdx < 2; ++wdx) {
      const int wlo = blk + 16 * wdx;
      if (16 * wdx < bn) {  // uniform
        // one pass per run the window meets (one for almost all windows),
        // the points outside the run zeroed: a single code path keeps every
        // accumulator in one register set
        const int whi = blk + min(16 * wdx + 16, bn);
        int p = wlo;
        do {
          const int vhi = min(whi, run_end);
          mfma_pass(buf, wdx, blk, true, p, vhi);
          p = vhi;
          if (p == run_end) finish_run(p);
          pin_acc();
        } while (p < whi);
      }
      // half of the next block's staging after each window
      if (next) {
        if (wdx == 0) {
          stage_a(buf ^ 1, 0, (kPerThr + 1) / 2);
          z_seed();
          z_steps(buf ^ 1, 0, kZO / 2);
        } else {
          stage_a(buf ^ 1, (kPerThr + 1) / 2, kPerThr);
          z_steps(buf ^ 1, kZO / 2, kZO);
        }
      }
      pin_acc();
    }
    // block blk + 2 kWBlk's h into the buffer block blk used (its A is
    // built), its records and weights into the staging registers; block
    // blk + 3 kWBlk's loads in flight during the next block
    if (zg < 4) sH[buf][zg][zp] = ph;
    take();
    load_stage(blk + 3 * kWBlk);
    __syncthreads();
  }
}

// The matrix engine's sorted weight rows: ws.wsf[s][pos] = w_s[sidx[pos]].
static int sort_weight_rows(Plan& pl, int64_t Np, const double* w, int64_t ldw, int nsets,
                            hipStream_t s) {
  Workspace& ws = pl.ws;
  if (nsets < 1 || nsets > 64 || nsets > ws.cap_sets || nsets > ws.cap_wsf_sets || !ws.wsf ||
      Np > ws.cap_set_points) {
    set_error("dsht: sorting 1..64 weight sets within the workspace");
    return DSHT_EINVAL;
  }
  const int* nvalid = ws.seg_start + pl.nseg;
  k_fill_u32<<<1, 64, 0, s>>>(ws.wmax, nsets, 0u);
  DSHT_LAUNCH_CHECK();
  if (Np == 0) return DSHT_OK;
  k_sorted_idx<<<(int)((Np + 255) / 256), 256, 0, s>>>(Np, nvalid, ws.rec, ws.sidx);
  DSHT_LAUNCH_CHECK();
  // wt (point-major copy) lives in the second half of the wsf allocation
  float* wt = ws.wsf + (int64_t)ws.cap_wsf_sets * ws.ldws;
  k_transpose_w<<<dim3((unsigned)((Np + kTile - 1) / kTile), (nsets + kTile - 1) / kTile), 256, 0,
                  s>>>(w, ldw, Np, nsets, wt);
  DSHT_LAUNCH_CHECK();
  // rows cover every position a staged block reads: (Np rounded up to 64) + 64 <= ldws
  const int64_t n = (Np + 63) / 64 * 64 + 64;
  const int64_t per = (int64_t)kTile * kGatherTiles;
  k_gather_sets<<<(unsigned)((n + per - 1) / per), 256, 0, s>>>(wt, nsets, nvalid, ws.sidx, n,
                                                                ws.wsf, ws.ldws, ws.wmax);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_scatter_sets(Plan& pl, int64_t Np, const double* w, int64_t ldw, int nsets,
                        hipStream_t s) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  Workspace& ws = pl.ws;
  if (nsets > ws.cap_part_sets) {
    set_error("dsht: the weight-set scatter takes 1..64 sets within the workspace");
    return DSHT_EINVAL;
  }
  int rc;
  if ((rc = sort_weight_rows(pl, Np, w, ldw, nsets, s))) return rc;
  const int nmy = (pl.Nl + kOrders - 1) / kOrders;
  const int64_t nwg = (int64_t)8 * ((nchunks + 7) / 8) * nmy;
  const int64_t stride = part_set_stride(pl);
  auto go = [&](auto kern) {
    kern<<<(unsigned)nwg, kThreads, 0, s>>>(ws.rec, ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                            ws.run_base, ws.wsf, ws.ldws, nsets, ws.wmax, nchunks,
                                            nmy, (float*)ws.part, stride, ws.run_seg);
  };
  // up to 32 sets: the set-side layout (4 k x 8, 16 or 32 sets per row
  // block, no padding to 32 sets, A shared by all orders); more: sets as
  // rows, h_k z formed per wavefront (measured faster from 33 sets on)
  if (nsets <= 32) {
    const int nmy2 = (pl.Nl + kWOrders - 1) / kWOrders;
    const int64_t nwg2 = (int64_t)8 * ((nchunks + 7) / 8) * nmy2;
    auto go2 = [&](auto kern) {
      kern<<<(unsigned)nwg2, 256, 0, s>>>(ws.rec, ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                          ws.run_base, ws.wsf, ws.ldws, nsets, ws.wmax, nchunks,
                                          nmy2, (float*)ws.part, stride, ws.run_seg);
    };
    if (nsets > 16) go2(k_scatter_wsets<4, 2, 4>);
    else if (nsets > 8) go2(k_scatter_wsets<2, 2, 4>);
    else go2(k_scatter_wsets<1, 2, 4>);
  } else {
    go(k_scatter_sets<2>);
  }
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
