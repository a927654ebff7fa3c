// Stage (b), part 2, spreading variant of the scatter (DSHT_FP32 plans,
// DSHT_SCATTER_SPREAD): the same sums as the direct scatters,
//   v_k(m, run) = sum_{p in run} h_k(p) e^{-i m phi_p},  m < Nl,  k = 0..3
// (the v's of sht/shared_interp_funcs.py:88-105), computed per run as a
// type-1 non-uniform FFT:
//   1. spread  g_k(x) = sum_p h_k(p) psi(x - t_p)  on the periodic grid
//      x = 0..n-1, t_p = n u_p = n phi_p / 2pi, with the exponential of
//      semicircle kernel psi(x) = exp(beta (sqrt(1 - (2x/W)^2) - 1)),
//      |x| <= W/2 (W taps);
//   2. transform  G_k(m) = sum_x g_k(x) e^{-2 pi i m x / n};
//   3. deconvolve  v_k(m) = G_k(m) / psihat(m),
//      psihat(m) = int psi(x) e^{-2 pi i m x / n} dx.
// The modes are one-sided (0 <= m < Nl) and n >= 4 Nl, so every wanted
// frequency is at most 1/4 per grid step and its aliases lie at 3/4 and
// beyond, where psihat has decayed by e^-beta-ish: W = 8, beta = 2.3 W gives
// ~2e-8 relative error, under the FP32 rounding of the grid.  The work per
// point is 4 W grid updates instead of 4 Nl products.
//
// One workgroup (1024 threads) per chunk of sorted points; the chunk's runs
// (stretches of one segment) are processed one after another in LDS.  The
// four real grids travel as two complex ones, g0 + i g1 and g2 + i g3
// ([2][n] float2), spread with LDS float atomics and transformed in place
// (radix-4 decimation in frequency, one radix-2 stage first when log2 n is
// odd; the transform is left digit-reversed and read where it lies), then
// split into the four real-data transforms, scaled by 1/psihat and written
// to part[run][k][m][2] (FP32), the layout the fold reads.
#include "common.h"
#include "plan.h"

namespace dsht {

constexpr int kSpreadThreads = 1024;

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-place DIF transform of the two grids sg[0..n) and sg[n..2n); tw[k] =
// e^{-2 pi i k / n}.  Frequency m ends up at position digit_rev(m).
__device__ inline void fft_pair(float2* sg, int n, int logn, const float2* __restrict__ tw) {
  const int tid = threadIdx.x;
  int S = n;
  if (logn & 1) {  // radix 2, span n
    const int q = n >> 1;
    for (int k = tid; k < n; k += kSpreadThreads) {  // n/2 butterflies per grid
      const int g = k >= q;
      const int j = k - g * q;
      float2* x = sg + g * n + j;
      const float2 a = x[0], b = x[q];
      x[0] = cadd(a, b);
      x[q] = cmulf(csub(a, b), tw[j]);
    }
    S = q;
    __syncthreads();
  }
  for (; S >= 4; S >>= 2) {  // radix 4, span S
    const int q = S >> 2, lq = __ffs(q) - 1;
    const int ts = n / S;
    const int nb = n >> 2;
    for (int k = tid; k < 2 * nb; k += kSpreadThreads) {  // n/4 butterflies per grid
      const int g = k >= nb;
      const int bi = k - g * nb;
      const int grp = bi >> lq, j = bi & (q - 1);
      float2* x = sg + g * n + grp * S + j;
      const float2 a = x[0], b = x[q], c = x[2 * q], d = x[3 * q];
      const float2 s0 = cadd(a, c), d0 = csub(a, c), s1 = cadd(b, d), d1 = csub(b, d);
      // y1 = (a - c) - i (b - d),  y3 = (a - c) + i (b - d)
      const float2 y1 = make_float2(d0.x + d1.y, d0.y - d1.x);
      const float2 y3 = make_float2(d0.x - d1.y, d0.y + d1.x);
      x[0] = cadd(s0, s1);
      x[q] = cmulf(y1, tw[j * ts]);
      x[2 * q] = cmulf(csub(s0, s1), tw[2 * j * ts]);
      x[3 * q] = cmulf(y3, tw[3 * j * ts]);
    }
    __syncthreads();
  }
}

// position of frequency m after fft_pair: its digits (lowest first, radix 2
// for the first stage when log2 n is odd, then radix 4) placed from the top
__device__ inline int digit_rev(int m, int logn) {
  int p = 0, span = 1 << logn;
  if (logn & 1) {
    span >>= 1;
    p += (m & 1) * span;
    m >>= 1;
  }
  while (span > 1) {
    span >>= 2;
    p += (m & 3) * span;
    m >>= 2;
  }
  return p;
}

// wmode: 0 the records' own weights, 1 w[sorted position], 2 w[rec.idx]
template <int kW>
__global__ __launch_bounds__(kSpreadThreads) void k_scatter_spread(
    const PointRec* __restrict__ rec, const int* __restrict__ seg_start, int nseg, int chunk,
    int Nl, const int* __restrict__ run_base, const double* __restrict__ w, int wmode, int logn,
    float beta, const float2* __restrict__ tw, const float* __restrict__ ipsi,
    float* __restrict__ part, int* __restrict__ run_seg) {
  extern __shared__ float2 sg[];  // [2][n]
  const int n = 1 << logn;
  const int tid = threadIdx.x;
  const int c = blockIdx.x;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;
  int run = run_base[c];
  for (int p = b; p < e;) {  // uniform: one run per iteration
    const int cur = rec[p].seg;
    const int pe = min(e, seg_start[cur + 1]);
    for (int i = tid; i < 2 * n; i += kSpreadThreads) sg[i] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int q = p + tid; q < pe; q += kSpreadThreads) {
      const PointRec& r = rec[q];
      const double wt = wmode == 0 ? 1.0 : (wmode == 1 ? w[q] : w[r.idx]);
      const float h0 = (float)(wt * r.h0), h1 = (float)(wt * r.h1);
      const float h2 = (float)(wt * r.h2), h3 = (float)(wt * r.h3);
      const double t = r.u * (double)n;
      const int j0 = (int)ceil(t - 0.5 * kW);
      const float z0 = (float)((double)j0 - t);  // in [-W/2, -W/2 + 1)
#pragma unroll
      for (int i = 0; i < kW; ++i) {
        const float z = (z0 + (float)i) * (2.0f / kW);
        const float psi = expf(beta * (sqrtf(fmaxf(0.f, 1.f - z * z)) - 1.f));
        const int x = (j0 + i) & (n - 1);
        atomicAdd(&sg[x].x, h0 * psi);
        atomicAdd(&sg[x].y, h1 * psi);
        atomicAdd(&sg[n + x].x, h2 * psi);
        atomicAdd(&sg[n + x].y, h3 * psi);
      }
    }
    __syncthreads();
    fft_pair(sg, n, logn, tw);
    // Z = G_a + i G_b for real grids a, b:  G_a(m) = (Z(m) + conj Z(n-m)) / 2,
    // G_b(m) = (Z(m) - conj Z(n-m)) / 2i
    for (int m = tid; m < Nl; m += kSpreadThreads) {
      const int pa = digit_rev(m, logn), pb = digit_rev((n - m) & (n - 1), logn);
      const float s = 0.5f * ipsi[m];
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const float2 z = sg[g * n + pa], v = sg[g * n + pb];
        const float2 ga = make_float2((z.x + v.x) * s, (z.y - v.y) * s);
        const float2 gb = make_float2((z.y + v.y) * s, (v.x - z.x) * s);
        *reinterpret_cast<float2*>(part + (((int64_t)run * 4 + 2 * g) * Nl + m) * 2) = ga;
        *reinterpret_cast<float2*>(part + (((int64_t)run * 4 + 2 * g + 1) * Nl + m) * 2) = gb;
      }
    }
    if (tid == 0) run_seg[run] = cur;
    ++run;
    p = pe;
    __syncthreads();  // the grids are read before the next run clears them
  }
}

int launch_scatter_spread(Plan& pl, int64_t Np, const double* w, int wmode, hipStream_t s) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  const int n = 1 << pl.spread_logn;
  const size_t lds = (size_t)2 * n * sizeof(float2);
  static bool attr = false;  // allow up to 128 KB of dynamic LDS (n = 8192)
  if (!attr) {
    DSHT_HIP(hipFuncSetAttribute((const void*)k_scatter_spread<kSpreadW>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 2 * 8192 * 8));
    attr = true;
  }
  k_scatter_spread<kSpreadW><<<nchunks, kSpreadThreads, lds, s>>>(
      pl.ws.rec, pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl, pl.ws.run_base, w, wmode,
      pl.spread_logn, (float)pl.spread_beta, pl.spread_tw, pl.spread_ipsi, (float*)pl.ws.part,
      pl.ws.run_seg);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
