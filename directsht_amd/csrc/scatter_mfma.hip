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
// One workgroup (4 wavefronts) per (chunk of sorted points, m tile of 1024
// orders).  Points are consumed in windows of 16 (one MFMA K step) that never
// cross a segment boundary; a shorter window at the end of a run is padded
// with h = 0.  Per window the workgroup generates E and B for its 16 points in
// FP64 (powers of e1 = e^{-i phi}) into LDS, B already split into the
// per-lane MFMA operand layout; wave w then builds the A operand of its rows
// (m1 in [8w, 8w+8), all k) and issues 12 MFMAs.  A run's result goes to its
// own slot part[run][half][m][4] (FP32), as for the vector scatter.
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

constexpr int kWin = 16;

// x = hi + lo with hi = fp16(x) and lo = fp16(x - hi).  x and hi are pinned
// in registers: otherwise the compiler may produce the stored hi from one
// rounding path (e.g. a packed conversion of the FP32 product) and lo from
// another (a mixed-precision FMA against a differently rounded hi), and lo
// then no longer corrects the hi that is actually used.
__device__ inline void split_f16(float x, _Float16& hi, _Float16& lo) {
  asm volatile("" : "+v"(x));
  _Float16 h = (_Float16)x;
  asm volatile("" : "+v"(h));  // one conversion: the stored hi is the subtracted hi
  hi = h;
  lo = (_Float16)(x - (float)h);
}

__global__ __launch_bounds__(256) void k_scatter_mfma(const PointRec* __restrict__ rec,
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

  __shared__ _Float16 sB[4][64][8];   // B operand per lane: re hi, re lo, im hi, im lo
  __shared__ float sE[2][32][kWin];   // E = e^{-i(1024 mt + 32 m1) phi_p}: [re|im][m1][p]
  __shared__ float sH[4][kWin];       // h_k(p), 0 for padding points

  // [row tile T: k in {2T, 2T+1}][column tile: re | im]; acc: MFMA
  // accumulator of the current group of windows, out: the run's outer sum
  constexpr int kGroup = 4;
  floatx16 acc[2][2], out[2][2];
  const floatx16 zero = {};
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int q = 0; q < 2; ++q) acc[T][q] = out[T][q] = zero;
  int wgroup = 0;  // windows accumulated in acc

  // this lane's A rows: i = l % 32 -> g, c1, m1; points 8*(l/32) + j
  const int ia = l & 31, bq = l >> 5;
  const int ga = ia >> 4, ca = (ia >> 3) & 1, m1a = 8 * w + (ia & 7);

  int run = run_base[c];
  int cur = rec[b].seg;
  int pos = b;
  while (pos < e) {
    const int run_end = min(e, seg_start[cur + 1]);
    while (pos < run_end) {
      const int nwin = min(kWin, run_end - pos);
      {
        // generation: thread (p = tid/16, q = tid%16) makes B for m2 = 2q, 2q+1
        // and E for m1 = 2q, 2q+1 of window point p
        const int p = tid >> 4, q = tid & 15;
        const bool valid = p < nwin;
        cplx e1{1.0, 0.0};
        double hk = 0.0;
        if (valid) {
          const PointRec& pr = rec[pos + p];
          e1 = cplx{pr.ere, pr.eim};
          if (q < 4) hk = q == 0 ? pr.h0 : q == 1 ? pr.h1 : q == 2 ? pr.h2 : pr.h3;
        }
        cplx z0 = cpow(e1, 2 * q);
        cplx z1 = cmul(z0, e1);
        cplx e32 = csq(csq(csq(csq(csq(e1)))));
        cplx E0 = cpow(e32, 32 * mt + 2 * q);
        cplx E1 = cmul(E0, e32);
        if (!valid) z0 = z1 = E0 = E1 = cplx{0.0, 0.0};
        const int lane0 = 32 * (p >> 3), j = p & 7;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const cplx z = u ? z1 : z0;
          const int ln = lane0 + 2 * q + u;
          split_f16((float)z.x, sB[0][ln][j], sB[1][ln][j]);
          split_f16((float)z.y, sB[2][ln][j], sB[3][ln][j]);
        }
        sE[0][2 * q][p] = (float)E0.x;
        sE[1][2 * q][p] = (float)E0.y;
        sE[0][2 * q + 1][p] = (float)E1.x;
        sE[1][2 * q + 1][p] = (float)E1.y;
        if (q < 4) sH[q][p] = (float)hk;
      }
      __syncthreads();
      {
        half8 ahi[2], alo[2];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const int p = 8 * bq + jj;
          const float ev = sE[ca][m1a][p];
#pragma unroll
          for (int T = 0; T < 2; ++T) {
            _Float16 hi, lo;
            split_f16(sH[2 * T + ga][p] * ev, hi, lo);
            ahi[T][jj] = hi;
            alo[T][jj] = lo;
          }
        }
        const half8 brh = *reinterpret_cast<const half8*>(sB[0][l]);
        const half8 brl = *reinterpret_cast<const half8*>(sB[1][l]);
        const half8 bih = *reinterpret_cast<const half8*>(sB[2][l]);
        const half8 bil = *reinterpret_cast<const half8*>(sB[3][l]);
        const bool fresh = wgroup == 0;  // uniform
#pragma unroll
        for (int T = 0; T < 2; ++T) {
          acc[T][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi[T], brh, fresh ? zero : acc[T][0], 0, 0, 0);
          acc[T][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi[T], brl, acc[T][0], 0, 0, 0);
          acc[T][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo[T], brh, acc[T][0], 0, 0, 0);
          acc[T][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi[T], bih, fresh ? zero : acc[T][1], 0, 0, 0);
          acc[T][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi[T], bil, acc[T][1], 0, 0, 0);
          acc[T][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo[T], bih, acc[T][1], 0, 0, 0);
        }
      }
      __syncthreads();
      pos += nwin;
      if (++wgroup == kGroup || pos >= run_end) {
#pragma unroll
        for (int T = 0; T < 2; ++T)
#pragma unroll
          for (int q = 0; q < 2; ++q) out[T][q] += acc[T][q];
        wgroup = 0;
      }
    }
    // flush the run: lane holds (k = 2T + g, m1 = 8w + 4(l/32) + d, m2 = l%32)
#pragma unroll
    for (int T = 0; T < 2; ++T)
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int m = 1024 * mt + 32 * (8 * w + 4 * bq + d) + (l & 31);
          if (m < Nl) {
            const float vre = out[T][0][8 * g + d] - out[T][1][8 * g + 4 + d];
            const float vim = out[T][1][8 * g + d] + out[T][0][8 * g + 4 + d];
            *reinterpret_cast<float2*>(part + (((int64_t)run * 2 + T) * Nl + m) * 4 + 2 * g) =
                make_float2(vre, vim);
          }
        }
    if (tid == 0 && mt == 0) run_seg[run] = cur;
#pragma unroll
    for (int T = 0; T < 2; ++T)
#pragma unroll
      for (int q = 0; q < 2; ++q) out[T][q] = zero;
    if (pos < e) {
      ++run;
      cur = rec[pos].seg;
    }
  }
}

int launch_scatter_mfma(Plan& pl, int64_t Np, hipStream_t s) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  dim3 grid(nchunks, (pl.Nl + 1023) / 1024);
  k_scatter_mfma<<<grid, 256, 0, s>>>(pl.ws.rec, pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl,
                                      pl.ws.run_base, (float*)pl.ws.part, pl.ws.run_seg);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
