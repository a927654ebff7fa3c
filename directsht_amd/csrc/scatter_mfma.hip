// Stage (b), part 2, matrix-core variant of the scatter (DSHT_FP32 plans):
//   v_k(m, seg) = sum_{p in seg} h_k(p) e^{-i m phi_p},  k = 0..3,
// the v's of sht/shared_interp_funcs.py:88-105, as a GEMM on the MFMA units.
//
// Factor m = 1024 mt + 32 m1 + m2 (mt = blockIdx.y, m1, m2 in [0, 32)):
//   e^{-i m phi} = E(m1) B(m2),  E = e^{-i (1024 mt + 32 m1) phi},  B = e^{-i m2 phi}
// so for one run (stretch of one segment) of points p
//   V_k(m) = sum_p [h_k(p) E_p(m1)] B_p(m2)
// is a complex GEMM with rows (k, m1) and columns m2; written as one real GEMM
// with rows (k, m1, re|im) = 256 and columns (re|im, m2) = 64:
//   V_re = C[re][re] - C[im][im],  V_im = C[re][im] + C[im][re].
// It runs as v_mfma_f32_32x32x16_f16 with each operand split into fp16
// hi + lo (x = hi + lo to ~2^-22 relative, or 2^-25 absolute where lo is
// subnormal, which the MFMA honours) and three products hi*hi + hi*lo + lo*hi
// accumulated in FP32.  The MFMA adds every product to its accumulator with
// an FP32 rounding, so a long run would accumulate ~3 roundings per point at
// the size of the running sum; the accumulator is therefore restarted from
// zero every 4 windows (64 points) and added into an outer FP32 sum.
//
// One workgroup (8 wavefronts) per (chunk of sorted points, m tile of 1024
// orders).  The chunk is staged 128 points at a time: the workgroup generates
// E and B for the block by short FP64 chains (powers of e1 = e^{-i phi} and of
// e1^32) into LDS, B split into fp16 hi/lo in [m2][point] order so every MFMA
// operand is one aligned 16-byte LDS read.  The block is then consumed in
// 16-aligned windows (one MFMA K step); a window that contains a segment
// boundary is issued once per run, with the points of other runs masked to
// zero in A.  Wave w owns row tile w (k = w/2, 16 values of m1) and issues 6
// MFMAs per window.  A run's result goes to its own slot part[run][half][m][4]
// (FP32), as for the vector scatter.
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

namespace dsht {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

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
constexpr int kBlock = 128; // points staged in LDS per generation phase
constexpr int kGroup = 4;   // windows per MFMA accumulator restart
// LDS row strides padded to 4 (mod 32) dwords: the 32 lanes of a half-wave
// read 32 different rows (m2 or m1) at the same column, and with this stride
// each group of 8 lanes covers all 32 banks with its 16-byte reads.
constexpr int kRowB = kBlock + 8;   // halves per B row (68 dwords)
constexpr int kRowE = kBlock + 4;   // floats per E row (132 dwords)

// x = hi + lo with hi = fp16(x) and lo = fp16(x - hi).  x and hi are pinned
// in registers: otherwise the compiler may produce the stored hi from one
// rounding path (e.g. a packed conversion of the FP32 product) and lo from
// another (a mixed-precision FMA against a differently rounded hi), and lo
// then no longer corrects the hi that is actually used.
__device__ inline void split_f16(float x, _Float16& hi, _Float16& lo) {
  asm volatile("" : "+v"(x));
  _Float16 h = (_Float16)x;
  asm volatile("" : "+v"(h));
  hi = h;
  lo = (_Float16)(x - (float)h);
}

__device__ inline uint32_t pack2(_Float16 a, _Float16 b) {
  return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

__global__ __launch_bounds__(512, 4) void k_scatter_mfma(const PointRec* __restrict__ rec,
                                                      const int* __restrict__ seg_start, int nseg,
                                                      int chunk, int Nl,
                                                      const int* __restrict__ run_base,
                                                      float* __restrict__ part,
                                                      int* __restrict__ run_seg) {
  const int c = blockIdx.x, mt = blockIdx.y;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;

  // B = e^{-i m2 phi} as fp16 hi/lo, [re|im][hi|lo][m2][point]; E FP32
  // [re|im][m1][point]; h_k FP32 [k][point]; segment of each point
  __shared__ __attribute__((aligned(16))) _Float16 sB[2][2][32][kRowB];
  __shared__ __attribute__((aligned(16))) float sE[2][32][kRowE];
  __shared__ __attribute__((aligned(16))) float sH[4][kBlock];
  __shared__ int sSeg[kBlock];

  // wave w owns row tile w: k = w/2, m1 = 16 (w&1) + 8g + e; rows i = 16g + 8c + e
  const int ia = l & 31, bq = l >> 5;
  const int kA = w >> 1;
  const int gA = ia >> 4, cA = (ia >> 3) & 1;
  const int m1A = 16 * (w & 1) + 8 * gA + (ia & 7);

  floatx16 acc[2], out[2];  // [column tile re | im]
  const floatx16 zero = {};
  acc[0] = acc[1] = out[0] = out[1] = zero;
  int wgroup = 0;

  int run = run_base[c];
  int cur = rec[b].seg;
  int run_end = min(e, seg_start[cur + 1]);

  for (int blk = b; blk < e; blk += kBlock) {
    const int bn = min(kBlock, e - blk);
    {
      // generation: thread (pair pp = tid & 63 -> points 2pp, 2pp+1; group
      // g = tid >> 6 -> m2 in [4g, 4g+4) and m1 in [4g, 4g+4)), FP64 chains
      const int pp = tid & 63, g = tid >> 6;
      cplx bz[2], ez[2], e1v[2], e32v[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int p = 2 * pp + u;
        cplx e1{1.0, 0.0};
        if (p < bn) {
          const PointRec& pr = rec[blk + p];
          e1 = cplx{pr.ere, pr.eim};
          if (g < 4) sH[g][p] = (float)(g == 0 ? pr.h0 : g == 1 ? pr.h1 : g == 2 ? pr.h2 : pr.h3);
          if (g == 4) sSeg[p] = pr.seg;
        } else {
          if (g < 4) sH[g][p] = 0.f;
          if (g == 4) sSeg[p] = -1;
        }
        const cplx e4 = csq(csq(e1));
        const cplx e32 = csq(csq(csq(e4)));
        const cplx e128 = csq(csq(e32));
        bz[u] = cpow(e4, g);                                           // e1^(4g)
        ez[u] = cmul(cpow(e128, g), cpow(csq(csq(csq(e128))), mt));    // e1^(128g + 1024mt)
        e1v[u] = e1;
        e32v[u] = e32;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = 4 * g + q;
        _Float16 rh0, rl0, ih0, il0, rh1, rl1, ih1, il1;
        split_f16((float)bz[0].x, rh0, rl0);
        split_f16((float)bz[0].y, ih0, il0);
        split_f16((float)bz[1].x, rh1, rl1);
        split_f16((float)bz[1].y, ih1, il1);
        *reinterpret_cast<uint32_t*>(&sB[0][0][m][2 * pp]) = pack2(rh0, rh1);
        *reinterpret_cast<uint32_t*>(&sB[0][1][m][2 * pp]) = pack2(rl0, rl1);
        *reinterpret_cast<uint32_t*>(&sB[1][0][m][2 * pp]) = pack2(ih0, ih1);
        *reinterpret_cast<uint32_t*>(&sB[1][1][m][2 * pp]) = pack2(il0, il1);
        *reinterpret_cast<float2*>(&sE[0][m][2 * pp]) = make_float2((float)ez[0].x, (float)ez[1].x);
        *reinterpret_cast<float2*>(&sE[1][m][2 * pp]) = make_float2((float)ez[0].y, (float)ez[1].y);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          bz[u] = cmul(bz[u], e1v[u]);
          ez[u] = cmul(ez[u], e32v[u]);
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
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const int pg = blk + po + jj;
          const float a = (pg >= p && pg < vhi) ? hv[jj] * ev[jj] : 0.f;
          _Float16 hi, lo;
          split_f16(a, hi, lo);
          ahi[jj] = hi;
          alo[jj] = lo;
        }
        const bool fresh = wgroup == 0;  // uniform
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, brh, fresh ? zero : acc[0], 0, 0, 0);
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, brl, acc[0], 0, 0, 0);
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, brh, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bih, fresh ? zero : acc[1], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bil, acc[1], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bih, acc[1], 0, 0, 0);
        p = vhi;
        const bool run_done = p == run_end;
        if (++wgroup == kGroup || run_done) {
          out[0] += acc[0];
          out[1] += acc[1];
          wgroup = 0;
        }
        if (run_done) {
          // flush: lane holds (k = kA, m1 = 16(w&1) + 8g + 4(l/32) + d, m2 = l%32)
#pragma unroll
          for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int d = 0; d < 4; ++d) {
              const int m = 1024 * mt + 32 * (16 * (w & 1) + 8 * g + 4 * bq + d) + (l & 31);
              if (m < Nl) {
                const float vre = out[0][8 * g + d] - out[1][8 * g + 4 + d];
                const float vim = out[1][8 * g + d] + out[0][8 * g + 4 + d];
                *reinterpret_cast<float2*>(part + (((int64_t)run * 2 + (kA >> 1)) * Nl + m) * 4 +
                                           2 * (kA & 1)) = make_float2(vre, vim);
              }
            }
          if (tid == 0 && mt == 0) run_seg[run] = cur;
          out[0] = out[1] = zero;
          if (p < e) {
            ++run;
            const int off = p - blk;  // point p may lie past this block
            cur = off < bn ? sSeg[off] : rec[p].seg;
            run_end = min(e, seg_start[cur + 1]);
          }
        }
      }
    }
    __syncthreads();
  }
}

int launch_scatter_mfma(Plan& pl, int64_t Np, hipStream_t s) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  dim3 grid(nchunks, (pl.Nl + 1023) / 1024);
  k_scatter_mfma<<<grid, 512, 0, s>>>(pl.ws.rec, pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                      pl.ws.run_base, (float*)pl.ws.part, pl.ws.run_seg);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
