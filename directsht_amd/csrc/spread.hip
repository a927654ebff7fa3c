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
// ~2e-8 relative error, under the FP32 rounding of the transform.  The work
// per point is 4 W grid updates instead of 4 Nl products.
//
// Two kernels: k_spread_nodes (grids of up to 4096 cells with 512 or more
// node pairs: single transforms, and the weight sets and dense point sets
// of batches, straight into their contraction slots) spreads per node and
// includes the node/parity fold (below, before its launcher);
// k_scatter_spread (the rest) spreads per run and leaves the fold to
// fold.hip.  Common to both:
//   * the kernel values 2^(beta log2(e) (sqrt(1 - z^2) - 1)) come from the
//     hardware sqrt and exp2 (es_kernel), and 1/psihat is the transform of
//     that very function (spread_kernel_values at plan creation).
//
// k_scatter_spread: one workgroup per chunk of sorted points; the chunk's
// runs (stretches of one segment) are processed one after another, each in
// LDS:
//   * the grids are accumulated in 64-bit fixed point (units of 2^-40 of the
//     weight scale) with LDS integer atomics.  Integer addition is
//     associative, so the grid -- and everything after it -- is the same bit
//     for bit whatever order the atomics land in; FP32 LDS atomics measured
//     7x slower on this part, and their order (hence their rounding) would
//     change from call to call;
//   * the real grids g0..g3 travel as complex grids Z = g0 + i g1 and
//     g2 + i g3 (or one of them per pass where n = 8192 leaves room for two
//     grids only), transformed by radix-8 decimation-in-frequency passes
//     (a radix-4 or radix-2 pass last when log2 n is not a multiple of 3),
//     one butterfly per thread per radix-8 pass, the first pass reading the
//     fixed-point grids and rounding them to FP32; the spectrum is left in
//     digit-reversed order and read where it lies;
//   * the four real-data transforms are split off Z(m) and Z(n - m), scaled
//     by 1/psihat(m) and the weight scale and written to part[run][k][m][2]
//     (FP32), the layout the fold reads.
#include "common.h"
#include "plan.h"

namespace dsht {

namespace {

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_mi(float2 a) { return make_float2(a.y, -a.x); }  // a * (-i)

// in-place DFT of size R (2, 4 or 8) of a[0..R): a[k] <- sum_t a[t] w_R^{kt},
// w_R = e^{-2 pi i / R}
template <int R>
__device__ __forceinline__ void dft(float2 (&a)[8]);
template <>
__device__ __forceinline__ void dft<2>(float2 (&a)[8]) {
  const float2 t = a[0];
  a[0] = c_add(t, a[1]);
  a[1] = c_sub(t, a[1]);
}
__device__ __forceinline__ void dft4(float2 x0, float2 x1, float2 x2, float2 x3, float2& X0,
                                     float2& X1, float2& X2, float2& X3) {
  const float2 s0 = c_add(x0, x2), d0 = c_sub(x0, x2);
  const float2 s1 = c_add(x1, x3), d1 = c_mi(c_sub(x1, x3));
  X0 = c_add(s0, s1);
  X2 = c_sub(s0, s1);
  X1 = c_add(d0, d1);
  X3 = c_sub(d0, d1);
}
template <>
__device__ __forceinline__ void dft<4>(float2 (&a)[8]) {
  dft4(a[0], a[1], a[2], a[3], a[0], a[1], a[2], a[3]);
}
template <>
__device__ __forceinline__ void dft<8>(float2 (&a)[8]) {
  // even outputs: DFT4 of u_t = a_t + a_{t+4}; odd: DFT4 of v_t = (a_t - a_{t+4}) w8^t
  constexpr float c = 0.70710678118654752f;
  const float2 u0 = c_add(a[0], a[4]), u1 = c_add(a[1], a[5]);
  const float2 u2 = c_add(a[2], a[6]), u3 = c_add(a[3], a[7]);
  const float2 v0 = c_sub(a[0], a[4]);
  const float2 e1 = c_sub(a[1], a[5]), e2 = c_sub(a[2], a[6]), e3 = c_sub(a[3], a[7]);
  const float2 v1 = make_float2(c * (e1.x + e1.y), c * (e1.y - e1.x));   // * (1 - i)/sqrt2
  const float2 v2 = c_mi(e2);                                            // * -i
  const float2 v3 = make_float2(c * (e3.y - e3.x), -c * (e3.x + e3.y));  // * (-1 - i)/sqrt2
  dft4(u0, u1, u2, u3, a[0], a[2], a[4], a[6]);
  dft4(v0, v1, v2, v3, a[1], a[3], a[5], a[7]);
}

// FFT work arrays: one float2 of padding after every 32 and another after
// every 64, which keeps every pass's and the spectrum read's LDS accesses
// within 3 bank cycles per wavefront (2 is the floor for 8-byte accesses)
__device__ __forceinline__ int fpad(int i) { return i + (i >> 5) + (i >> 6); }

// 64-bit fixed point: round(a b) for |a b| < 2^51, by one FMA onto the FP64
// shift 1.5 * 2^52 (exact product, one rounding to an integer); the shift's
// bits are 0x4338000000000000, whose low word is zero.
__device__ __forceinline__ unsigned long long to_fixed(double a, double b) {
  const double d = __builtin_fma(a, b, 0x1.8p52);
  return (unsigned long long)(__double_as_longlong(d) - 0x4338000000000000LL);
}

// v 2^-40 in FP32 for a 64-bit fixed-point sum |v| < 2^54: the high word
// (|hi| < 2^22) converts exactly, so the only roundings are the low word's
// (below 2^-32 absolute) and the final FMA's
__device__ __forceinline__ float from_fixed(unsigned long long v) {
  const int hi = (int)(v >> 32);
  const unsigned lo = (unsigned)v;
  return __builtin_fmaf((float)hi, 0x1p-8f, (float)lo * 0x1p-40f);
}

// The kernel at z = (x - t) 2 / W, z in [-1, 1]:
//   psi = exp(beta (sqrt(1 - z^2) - 1)) = 2^(beta log2(e) (sqrt(1 - z^2) - 1)),
// with the hardware square root and base-2 exponential (about 1 ulp each; the
// exponent lies in [-beta log2(e), 0], so neither needs range handling):
// six instructions per tap, where the correctly rounded sqrtf and expf
// expand to about thirty.  Their rounding is not centred (it scaled the
// transform by ~2e-7), so the deconvolution divides by the transform of
// this very function: its values at the quadrature nodes come from the
// device (spread_kernel_values).
__device__ __forceinline__ float es_kernel(float z, float beta_log2e) {
  const float s = __builtin_amdgcn_sqrtf(fmaxf(0.f, __builtin_fmaf(-z, z, 1.f)));
  return __builtin_amdgcn_exp2f(__builtin_fmaf(beta_log2e, s, -beta_log2e));
}

}  // namespace

// Grid geometry of the kernel for n = 2^LOGN: G real grids held per pass over
// the run's points (4, or 2 where n = 8192), NCG = G / 2 complex grids, one
// radix-8 butterfly per thread.
template <int LOGN>
struct SpreadGeom {
  static constexpr int n = 1 << LOGN;
  static constexpr int G = LOGN >= 13 ? 2 : 4;
  static constexpr int NCG = G / 2;
  static constexpr int kPasses8 = LOGN / 3;           // radix-8 passes
  static constexpr int kRadixLast = 1 << (LOGN % 3);  // then one pass of 2 or 4 (1: none)
  static constexpr int kThreads = NCG * n / 8;
  static constexpr int kPad = n + n / 32 + n / 64;    // complex grid stride (float2)
  // radix-8 passes with twiddles (span above 8), and how many of them keep
  // each thread's twiddles in registers for the whole kernel (the butterfly
  // a thread takes in a pass is the same in every run)
  static constexpr int kTwPasses = kPasses8 - (kRadixLast == 1 ? 1 : 0);
  static constexpr int kTwReg = kTwPasses < 3 ? kTwPasses : 3;
  static_assert(kThreads >= 64 && kThreads <= 1024, "unsupported grid size");
  static_assert((int64_t)G * n * 8 <= 128 * 1024, "grids exceed the LDS budget");
  static_assert(NCG * kPad <= G * n, "the FFT's grids must fit in the spread grids' LDS");
};

// position of frequency f in the digit-reversed spectrum
template <int LOGN>
__device__ __forceinline__ int spec_pos(int f) {
  using Gm = SpreadGeom<LOGN>;
  int p = 0, span = Gm::n;
#pragma unroll
  for (int s = 0; s < Gm::kPasses8; ++s) {
    span >>= 3;
    p += (f & 7) * span;
    f >>= 3;
  }
  if (Gm::kRadixLast > 1) p += f & (Gm::kRadixLast - 1);  // span 1
  return p;
}

// One in-place DIF pass of radix R and span S over the NCG complex grids
// (butterflies bi = tid, tid + kThreads, ...).  The twiddles of radix-8 pass
// P come from twr[P] (this thread's, loaded once) for P < TWR (at most
// kTwReg), else from the table tw[k] = e^{-2 pi i k / n}.
template <int LOGN, int R, int S, int P, int TWR = SpreadGeom<LOGN>::kTwReg>
__device__ __forceinline__ void fft_pass(float2* __restrict__ sz, const float2 (&twr)[3][7],
                                         const float2* __restrict__ tw, int tid) {
  using Gm = SpreadGeom<LOGN>;
  constexpr int n = Gm::n, q = S / R, per = Gm::NCG * n / R / Gm::kThreads, ts = n / S;
#pragma unroll
  for (int k = 0; k < per; ++k) {
    const int bi = tid + k * Gm::kThreads;
    const int c = bi / (n / R), b = bi % (n / R);
    const int g = b / q, j = b % q;
    float2* base = sz + c * Gm::kPad;
    const int i0 = g * S + j;
    float2 a[8];
#pragma unroll
    for (int t = 0; t < R; ++t) a[t] = base[fpad(i0 + t * q)];
    dft<R>(a);
    if (S > R) {
#pragma unroll
      for (int t = 1; t < R; ++t)
        a[t] = c_mul(a[t], P < TWR ? twr[P < 3 ? P : 0][t - 1] : tw[t * j * ts]);
    }
#pragma unroll
    for (int t = 0; t < R; ++t) base[fpad(i0 + t * q)] = a[t];
  }
}

template <int LOGN, int P, int TWR = SpreadGeom<LOGN>::kTwReg>
__device__ __forceinline__ void fft_passes8(float2* sz, const float2 (&twr)[3][7],
                                            const float2* tw, int tid) {
  if constexpr (P < SpreadGeom<LOGN>::kPasses8) {
    fft_pass<LOGN, 8, (SpreadGeom<LOGN>::n >> (3 * P)), P, TWR>(sz, twr, tw, tid);
    __syncthreads();
    fft_passes8<LOGN, P + 1, TWR>(sz, twr, tw, tid);
  }
}

// wmode 0: the records' own weights, 1: w[sorted position], 2: w[rec.idx]
template <int LOGN>
__global__ __launch_bounds__(SpreadGeom<LOGN>::kThreads) void k_scatter_spread(
    const SpreadRec* __restrict__ rec, const int* __restrict__ seg_start, int nseg, int chunk,
    int Nl, const int* __restrict__ run_base, const double* __restrict__ w, int wmode,
    float beta_log2e, const float2* __restrict__ tw, const float* __restrict__ ipsi,
    float* __restrict__ part, int* __restrict__ run_seg, const unsigned* __restrict__ wmax) {
  using Gm = SpreadGeom<LOGN>;
  constexpr int n = Gm::n, G = Gm::G, NCG = Gm::NCG, kT = Gm::kThreads;
  constexpr int kW = kSpreadW;
  __shared__ __attribute__((aligned(16))) unsigned long long sgrid[G * n];
  float2* sz = reinterpret_cast<float2*>(sgrid);  // the FFT's grids, after the spread
  const int tid = threadIdx.x;
  const int c = blockIdx.x;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) return;
  // this thread's twiddles of the first kTwReg radix-8 passes: pass P has
  // span S = n / 8^P and takes butterfly j = (tid mod n/8) mod S/8
  float2 twr[3][7];
#pragma unroll
  for (int P = 0; P < 3; ++P) {
    if (P < Gm::kTwReg) {
      const int S = n >> (3 * P), j = (tid % (n / 8)) % (S / 8);
#pragma unroll
      for (int t = 1; t < 8; ++t) twr[P][t - 1] = tw[t * j * (n / S)];
    }
  }

  // weight scale: 2^-ex brings max |w| below 1, so every fixed-point
  // contribution stays below 2^40 and a cell's sum below 2^54; undone exactly
  // at the output together with the 2^-40 of the fixed point
  int ex = 0;
  {
    const float mx = __uint_as_float(wmax[0]);
    if (mx > 0.f && mx <= 3.0e38f) (void)frexpf(mx, &ex);
    ex = max(-120, min(120, ex));
  }
  const double w_fix = ldexp(1.0, 40 - ex);
  const float w_out = ldexpf(1.0f, ex);

  int run = run_base[c];
  for (int p = b; p < e;) {  // uniform: one run per iteration
    const int cur = rec[p].seg;  // uniform
    const int pe = min(e, seg_start[cur + 1]);
#pragma unroll 1
    for (int pass = 0; pass < 4 / G; ++pass) {
      for (int i = tid; i < G * n / 2; i += kT)
        reinterpret_cast<ulonglong2*>(sgrid)[i] = make_ulonglong2(0ull, 0ull);
      __syncthreads();
      for (int q = p + tid; q < pe; q += kT) {
        const float4 ah = reinterpret_cast<const float4*>(rec + q)[0];
        const int4 ri = reinterpret_cast<const int4*>(rec + q)[1];  // seg, idx, j0, z0
        const double wt = (wmode == 0 ? 1.0 : (wmode == 1 ? w[q] : w[ri.y])) * w_fix;
        const float hk[4] = {ah.x, ah.y, ah.z, ah.w};
        double a[G];
#pragma unroll
        for (int g = 0; g < G; ++g) a[g] = wt * (double)hk[pass * G + g];
        const int j0 = ri.z;
        const float zs = __int_as_float(ri.w) * (2.0f / kW);  // z0 in [-W/2, -W/2 + 1), scaled
#pragma unroll
        for (int i = 0; i < kW; ++i) {
          const float z = zs + (float)i * (2.0f / kW);
          const double psi = (double)es_kernel(z, beta_log2e);
          const int x = (j0 + i) & (n - 1);
#pragma unroll
          for (int g = 0; g < G; ++g) atomicAdd(&sgrid[g * n + x], to_fixed(a[g], psi));
        }
      }
      __syncthreads();

      // first radix-8 pass straight from the fixed-point grids (span n)
      {
        constexpr int q = n / 8;
        const int cg = tid / q, j = tid % q;
        float2 v[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int x = j + t * q;
          v[t] = make_float2(from_fixed(sgrid[(2 * cg) * n + x]),
                             from_fixed(sgrid[(2 * cg + 1) * n + x]));
        }
        __syncthreads();  // every fixed-point value is read before sz overwrites it
        dft<8>(v);
#pragma unroll
        for (int t = 1; t < 8; ++t) v[t] = c_mul(v[t], twr[0][t - 1]);
        float2* base = sz + cg * Gm::kPad;
#pragma unroll
        for (int t = 0; t < 8; ++t) base[fpad(j + t * q)] = v[t];
      }
      __syncthreads();
      fft_passes8<LOGN, 1>(sz, twr, tw, tid);
      if constexpr (Gm::kRadixLast > 1) {
        fft_pass<LOGN, Gm::kRadixLast, Gm::kRadixLast, Gm::kPasses8>(sz, twr, tw, tid);
        __syncthreads();
      }

      // Z = G_a + i G_b for real grids a, b:  G_a(m) = (Z(m) + conj Z(n-m)) / 2,
      // G_b(m) = (Z(m) - conj Z(n-m)) / 2i
      for (int m = tid; m < Nl; m += kT) {
        const int pa = fpad(spec_pos<LOGN>(m)), pb = fpad(spec_pos<LOGN>((n - m) & (n - 1)));
        const float s = 0.5f * ipsi[m] * w_out;
#pragma unroll
        for (int cg = 0; cg < NCG; ++cg) {
          const float2 z = sz[cg * Gm::kPad + pa], v = sz[cg * Gm::kPad + pb];
          const float2 ga = make_float2((z.x + v.x) * s, (z.y - v.y) * s);
          const float2 gb = make_float2((z.y + v.y) * s, (v.x - z.x) * s);
          const int k = pass * G + 2 * cg;
          *reinterpret_cast<float2*>(part + (((int64_t)run * 4 + k) * Nl + m) * 2) = ga;
          *reinterpret_cast<float2*>(part + (((int64_t)run * 4 + k + 1) * Nl + m) * 2) = gb;
        }
      }
      __syncthreads();  // the spectrum is read before the next grids are cleared
    }
    if (tid == 0) run_seg[run] = cur;
    ++run;
    p = pe;
  }
}

// Node variant (plain transforms): the spreading scatter and the node/parity
// fold in one kernel, so the per-run partials never leave the workgroup.
// Workgroup = node pair np; its four units (node n = 2 np + odd, branch br)
// are done one after another.  Unit (n, br) spreads the points of bins n - 1
// and n of segments (br, h, .) onto four real grids, for h = 0, 1:
//   c^h = sum_{bin n} a0 psi + sum_{bin n-1} a2 psi,
//   d^h = sum_{bin n} a1 psi + sum_{bin n-1} a3 psi,
// whose deconvolved spectra are the fold's c and d of node n (the v0(n) +
// v2(n-1) of sht/shared_interp_funcs.py:88-105 summed per node, sht/sht.py:
// 266-277).  Each point is spread twice (once for each node it touches), two
// of its four weights each time: as many grid updates as the run-based
// kernel, but one FFT of four real grids per (node, branch) instead of one
// per (segment) piece: half the FFTs, and no partials written or read.  The
// thread of mode m keeps the first node's c^h, d^h at m in registers; at the
// second it forms the parity sums c_P = c^0 +- c^1, d_P (times dx) and the
// derivative coefficients f_P = x d_P / (1 - x^2), g = d_{1-P} / (1 - x^2)
// of the pair's two nodes and writes the branch's records (the fold's
// layout, plain_rec: [c | f | g][re, im][node], one contiguous block per
// pair and branch, whole lines from one workgroup).
// (One workgroup per node pair and branch measured slower.)  Grids up to
// n = 4096 (all four in LDS at once, one mode per thread).
// kSets: ns.count sets of a batch for the matrix-core contraction instead,
// set s into slot set0 + s (the 16-value rows rhs[m][P][np][slot of
// ntot][16] of the batched fold: branch br at 8 br, re then im, each
// c(2np), c(2np+1), dx d(2np), dx d(2np+1)), with the set's max(|c|, |d|)
// into rmax[s] (float bits, cleared beforehand).  Set s weighs its points by
// wts[input index * count + s] (FP32 weight sets, point-major, scaled by
// wmax[s]), or by the records' own weights if wts is NULL (wmax[0]).  The workgroups of one node pair for all sets run together on
// one XCD (block b -> XCD b mod 8, the order in which the hardware hands them
// out): they read the same records and weight rows (from that XCD's L2) and
// write the same coefficient rows at about the same time, so the sets'
// 32-byte pieces of a row meet in the caches and reach HBM as whole lines.
struct NodeSets {
  const float* wts;
  unsigned* rmax;
  int count, set0, ntot;
};
constexpr int kNodeBatch = 2;
template <int LOGN, bool kSets>
__global__ __launch_bounds__(SpreadGeom<LOGN>::kThreads) void k_spread_nodes(
    const SpreadRec* __restrict__ rec, const int* __restrict__ seg_start, int Nx, int Nl,
    int Npair, double dx, const double* __restrict__ w, int wmode, float beta_log2e,
    const float2* __restrict__ tw, const float* __restrict__ ipsi, float* __restrict__ rhs,
    const unsigned* __restrict__ wmax, NodeSets ns) {
  using Gm = SpreadGeom<LOGN>;
  constexpr int n = Gm::n, G = Gm::G, NCG = Gm::NCG, kT = Gm::kThreads;
  constexpr int kW = kSpreadW;
  static_assert(G == 4 && kT >= n / 4, "one pass over four grids, one mode per thread");
  __shared__ __attribute__((aligned(16))) unsigned long long sgrid[G * n];
  float2* sz = reinterpret_cast<float2*>(sgrid);
  const int tid = threadIdx.x;
  int np = blockIdx.x, sl = 0;  // node pair, set of the launch
  if constexpr (kSets) {  // XCD x: node pairs x, x + 8, ..., each for every set
    const int k = blockIdx.x >> 3;
    np = (blockIdx.x & 7) + 8 * (k / ns.count);
    sl = k % ns.count;
    if (np >= Npair) return;  // the padding of the last round
  }
  const float* wsf = kSets && ns.wts ? ns.wts + sl : nullptr;  // set sl's weights, stride count
  const int m = tid;  // this thread's mode
  // twiddles of the first radix-8 pass in registers (the spectra held across
  // the units leave no room for more), the others from the table
  constexpr int kTwr = Gm::kTwReg < 1 ? Gm::kTwReg : 1;
  float2 twr[3][7];
  if (kTwr > 0) {
    const int j = tid % (n / 8);
#pragma unroll
    for (int t = 1; t < 8; ++t) twr[0][t - 1] = tw[t * j];
  }
  int ex = 0;
  {
    const float mx = __uint_as_float(wmax[wsf ? sl : 0]);
    if (mx > 0.f && mx <= 3.0e38f) (void)frexpf(mx, &ex);
    ex = max(-120, min(120, ex));
  }
  const double w_fix = ldexp(1.0, 40 - ex);
  const float w_out = ldexpf(1.0f, ex);
  double xs[2], inv[2];
#pragma unroll
  for (int odd = 0; odd < 2; ++odd) {
    const int nd = 2 * np + odd;
    xs[odd] = (double)nd * dx;
    inv[odd] = nd < Nx ? 1.0 / (1.0 - xs[odd] * xs[odd]) : 0.0;
  }

  // the records of branch br at mode m from the pair's spectra
  // a[odd][h][c re, c im, d re, d im], as the fold writes them: per P,
  // [c | f | g] x (re, im) x (node 0, 1), 48 contiguous bytes (kSets: the
  // branch's 32 bytes of the set's row)
  float vmax = 0.f;
  auto emit = [&](int br, const float (&a0)[2][4], const float (&a1)[2][4]) {
#pragma unroll
    for (int P = 0; P < 2; ++P) {
      const double sg = P ? -1.0 : 1.0;
      float c4[4], f4[4], g4[4];
#pragma unroll
      for (int ri = 0; ri < 2; ++ri)
#pragma unroll
        for (int odd = 0; odd < 2; ++odd) {
          const float (&a)[2][4] = odd ? a1 : a0;
          const double cv = (double)a[0][ri] + sg * (double)a[1][ri];
          const double dv = dx * ((double)a[0][2 + ri] + sg * (double)a[1][2 + ri]);
          if constexpr (kSets) {
            c4[2 * ri + odd] = (float)cv;
            f4[2 * ri + odd] = (float)dv;
            vmax = fmaxf(vmax, (float)fmax(fabs(cv), fabs(dv / dx)));
          } else {
            const double du = dx * ((double)a[0][2 + ri] - sg * (double)a[1][2 + ri]);  // d_{1-P}
            c4[2 * ri + odd] = (float)cv;
            f4[2 * ri + odd] = (float)(xs[odd] * inv[odd] * dv);
            g4[2 * ri + odd] = (float)(inv[odd] * du);
          }
        }
      if constexpr (kSets) {
        float* o = rhs + ((((int64_t)m * 2 + P) * Npair + np) * ns.ntot + ns.set0 + sl) * 16 + 8 * br;
        *reinterpret_cast<float4*>(o) = make_float4(c4[0], c4[1], f4[0], f4[1]);
        *reinterpret_cast<float4*>(o + 4) = make_float4(c4[2], c4[3], f4[2], f4[3]);
      } else {
        float* o = rhs + plain_rec(np, br, m, P, Nl);
        *reinterpret_cast<float4*>(o) = make_float4(c4[0], c4[1], c4[2], c4[3]);
        *reinterpret_cast<float4*>(o + 4) = make_float4(f4[0], f4[1], f4[2], f4[3]);
        *reinterpret_cast<float4*>(o + 8) = make_float4(g4[0], g4[1], g4[2], g4[3]);
      }
    }
  };

#pragma unroll 1
  for (int br = 0; br < 2; ++br) {
    float a0[2][4];  // the pair's first node at mode m: [h][c re, c im, d re, d im]
#pragma unroll
    for (int odd = 0; odd < 2; ++odd) {
      const int node = 2 * np + odd;
      // hemisphere h: points [lo, hi) of bins node - 1 (below mid) and node
      int lo[2], mid[2], hi[2];
      bool any = false;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int base = (br * 2 + h) * (Nx - 1);
        const int jlo = max(node - 1, 0), jhi = min(node, Nx - 2);
        if (jlo <= jhi) {
          lo[h] = seg_start[base + jlo];
          hi[h] = seg_start[base + jhi + 1];
          mid[h] = node <= Nx - 2 ? seg_start[base + node] : hi[h];
        } else {
          lo[h] = mid[h] = hi[h] = 0;
        }
        any |= lo[h] < hi[h];
      }
      float a[2][4] = {};
      if (any) {  // uniform; no points at this node: zero spectra, no transform
        // a grid here takes all the node's points of one hemisphere (the
        // run-based kernel at most a work item's): 2^-sh more fixed-point
        // scale for more than 2^14 of them keeps every cell sum below 2^54
        const int cnt = max(hi[0] - lo[0], hi[1] - lo[1]);
        const int sh = max(0, 32 - __builtin_clz((unsigned)cnt) - 14);
        const double wf = ldexp(w_fix, -sh);
        const float wo = ldexpf(w_out, sh);
        for (int i = tid; i < G * n / 2; i += kT)
          reinterpret_cast<ulonglong2*>(sgrid)[i] = make_ulonglong2(0ull, 0ull);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          unsigned long long* gc = sgrid + (2 * h) * n;  // c^h, then d^h
          unsigned long long* gd = gc + n;
          // kNodeBatch points per thread and step: their record loads, then
          // their weight loads, are in flight together
          for (int q0 = lo[h] + tid; q0 < hi[h]; q0 += kNodeBatch * kT) {
            float4 ah[kNodeBatch];
            int4 ri[kNodeBatch];  // seg, idx, j0, z0
            double wt[kNodeBatch];
#pragma unroll
            for (int b = 0; b < kNodeBatch; ++b) {
              const int q = min(q0 + b * kT, hi[h] - 1);
              ah[b] = reinterpret_cast<const float4*>(rec + q)[0];
              ri[b] = reinterpret_cast<const int4*>(rec + q)[1];
            }
#pragma unroll
            for (int b = 0; b < kNodeBatch; ++b) {
              const int q = min(q0 + b * kT, hi[h] - 1);
              if (kSets)
                wt[b] = (wsf ? (double)wsf[(int64_t)ri[b].y * ns.count] : 1.0) * wf;
              else
                wt[b] = (wmode == 0 ? 1.0 : (wmode == 1 ? w[q] : w[ri[b].y])) * wf;
            }
#pragma unroll
            for (int b = 0; b < kNodeBatch; ++b) {
              const int q = q0 + b * kT;
              if (q >= hi[h]) break;
              const bool upper = q >= mid[h];  // bin node: weights a0, a1; bin node - 1: a2, a3
              const double ac = wt[b] * (double)(upper ? ah[b].x : ah[b].z);
              const double ad = wt[b] * (double)(upper ? ah[b].y : ah[b].w);
              const int j0 = ri[b].z;
              const float zs = __int_as_float(ri[b].w) * (2.0f / kW);  // z0, scaled
#pragma unroll
              for (int i = 0; i < kW; ++i) {
                const float z = zs + (float)i * (2.0f / kW);
                const double psi = (double)es_kernel(z, beta_log2e);
                const int x = (j0 + i) & (n - 1);
                atomicAdd(&gc[x], to_fixed(ac, psi));
                atomicAdd(&gd[x], to_fixed(ad, psi));
              }
            }
          }
        }
        __syncthreads();
        {
          constexpr int q = n / 8;
          const int cg = tid / q, j = tid % q;
          float2 v[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            const int x = j + t * q;
            v[t] = make_float2(from_fixed(sgrid[(2 * cg) * n + x]),
                               from_fixed(sgrid[(2 * cg + 1) * n + x]));
          }
          __syncthreads();
          dft<8>(v);
#pragma unroll
          for (int t = 1; t < 8; ++t) v[t] = c_mul(v[t], kTwr > 0 ? twr[0][t - 1] : tw[t * j]);
          float2* base = sz + cg * Gm::kPad;
#pragma unroll
          for (int t = 0; t < 8; ++t) base[fpad(j + t * q)] = v[t];
        }
        __syncthreads();
        fft_passes8<LOGN, 1, kTwr>(sz, twr, tw, tid);
        if constexpr (Gm::kRadixLast > 1) {
          fft_pass<LOGN, Gm::kRadixLast, Gm::kRadixLast, Gm::kPasses8, kTwr>(sz, twr, tw, tid);
          __syncthreads();
        }
        // complex grid h = c^h + i d^h: split and deconvolve
        if (m < Nl) {
          const int pa = fpad(spec_pos<LOGN>(m)), pb = fpad(spec_pos<LOGN>((n - m) & (n - 1)));
          const float s = 0.5f * ipsi[m] * wo;
#pragma unroll
          for (int h = 0; h < NCG; ++h) {
            const float2 z = sz[h * Gm::kPad + pa], v = sz[h * Gm::kPad + pb];
            a[h][0] = (z.x + v.x) * s;
            a[h][1] = (z.y - v.y) * s;
            a[h][2] = (z.y + v.y) * s;
            a[h][3] = (v.x - z.x) * s;
          }
        }
        __syncthreads();  // the spectrum is read before the next grids are cleared
      }
      if (odd == 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int q = 0; q < 4; ++q) a0[h][q] = a[h][q];
      } else if (m < Nl) {
        emit(br, a0, a);
      }
    }
  }
  if constexpr (kSets) {  // the set's bound: wavefront, then one atomic per wavefront
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
    if ((tid & 63) == 0) atomic_max_pos(ns.rmax + sl, vmax);
  }
}

__global__ void k_es_values(const float* __restrict__ z, int nq, float beta_log2e,
                            float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) out[i] = es_kernel(z[i], beta_log2e);
}

int spread_kernel_values(const float* z, int nq, double beta, float* out) {
  float* d = nullptr;
  DSHT_HIP(hipMalloc(&d, 2 * (size_t)nq * sizeof(float)));
  hipError_t e = hipMemcpy(d, z, nq * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    k_es_values<<<(nq + 255) / 256, 256>>>(d, nq, (float)(beta * M_LOG2E), d + nq);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d + nq, nq * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) {
    set_error(std::string("spread_kernel_values: ") + hipGetErrorString(e));
    return DSHT_EHIP;
  }
  return DSHT_OK;
}

template <int LOGN>
static int launch_nodes_n(Plan& pl, const double* w, int wmode, const unsigned* wmax,
                          const NodeSets* ns, hipStream_t s) {
  using Gm = SpreadGeom<LOGN>;
  const float bl = (float)(pl.spread_beta * M_LOG2E);
  const auto* rec = reinterpret_cast<const SpreadRec*>(pl.ws.rec);
  if (ns) {
    const int blocks = 8 * ((pl.Npair + 7) / 8) * ns->count;  // rounds of 8 node pairs x sets
    k_spread_nodes<LOGN, true><<<blocks, Gm::kThreads, 0, s>>>(
        rec, pl.ws.seg_start, pl.Nx, pl.Nl, pl.Npair, pl.dx, w, wmode, bl, pl.spread_tw,
        pl.spread_ipsi, (float*)pl.ws.rhs, wmax, *ns);
  } else {
    k_spread_nodes<LOGN, false><<<pl.Npair, Gm::kThreads, 0, s>>>(
        rec, pl.ws.seg_start, pl.Nx, pl.Nl, pl.Npair, pl.dx, w, wmode, bl, pl.spread_tw,
        pl.spread_ipsi, (float*)pl.ws.rhs, wmax, NodeSets{});
  }
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

static int nodes_dispatch(Plan& pl, const double* w, int wmode, const unsigned* wmax,
                          const NodeSets* ns, hipStream_t s) {
  switch (pl.spread_logn) {
    case 8: return launch_nodes_n<8>(pl, w, wmode, wmax, ns, s);
    case 9: return launch_nodes_n<9>(pl, w, wmode, wmax, ns, s);
    case 10: return launch_nodes_n<10>(pl, w, wmode, wmax, ns, s);
    case 11: return launch_nodes_n<11>(pl, w, wmode, wmax, ns, s);
    case 12: return launch_nodes_n<12>(pl, w, wmode, wmax, ns, s);
    default:
      set_error("dsht: the node spreading scatter supports grids up to 4096 cells");
      return DSHT_EINVAL;
  }
}

int launch_spread_nodes(Plan& pl, int64_t Np, const double* w, int wmode, hipStream_t s) {
  int rc;
  if (wmode != 0 && Np > 0 && (rc = launch_weight_max(pl, Np, w, s))) return rc;
  return nodes_dispatch(pl, w, wmode, pl.ws.wmax, nullptr, s);
}

int launch_spread_nodes_sets(Plan& pl, int set0, int count, int ntot, bool weight_rows,
                             hipStream_t s) {
  Workspace& ws = pl.ws;
  if (count < 1 || set0 < 0 || set0 + count > ntot || ntot > ws.cap_sets ||
      (int64_t)pl.Nl * 2 * pl.Npair * 16 * ntot > ws.rhs_elems ||
      (weight_rows && (count > ws.cap_wsf_sets || !ws.wsf))) {
    set_error("dsht: node spreading into batch slots beyond the workspace's sets");
    return DSHT_EINVAL;
  }
  NodeSets ns{weight_rows ? ws.wsf : nullptr, ws.rmax + set0, count, set0, ntot};
  int rc;
  if ((rc = launch_clear_u32(ws.rmax + set0, nullptr, count, s))) return rc;
  return nodes_dispatch(pl, nullptr, 0, ws.wmax, &ns, s);
}

template <int LOGN>
static int launch_spread_n(Plan& pl, int nchunks, const double* w, int wmode, const unsigned* wmax,
                           float* part, hipStream_t s) {
  using Gm = SpreadGeom<LOGN>;
  k_scatter_spread<LOGN><<<nchunks, Gm::kThreads, 0, s>>>(
      reinterpret_cast<const SpreadRec*>(pl.ws.rec), pl.ws.seg_start, pl.nseg, pl.chunk, pl.Nl, pl.ws.run_base, w, wmode,
      (float)(pl.spread_beta * M_LOG2E), pl.spread_tw, pl.spread_ipsi, part, pl.ws.run_seg, wmax);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_scatter_spread(Plan& pl, int64_t Np, const double* w, int wmode, hipStream_t s,
                          float* part, const unsigned* wmax) {
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks == 0) return DSHT_OK;
  int rc;
  if (!wmax) {
    if (wmode != 0 && (rc = launch_weight_max(pl, Np, w, s))) return rc;
    wmax = pl.ws.wmax;
  }
  if (!part) part = (float*)pl.ws.part;
  switch (pl.spread_logn) {
    case 8: return launch_spread_n<8>(pl, nchunks, w, wmode, wmax, part, s);
    case 9: return launch_spread_n<9>(pl, nchunks, w, wmode, wmax, part, s);
    case 10: return launch_spread_n<10>(pl, nchunks, w, wmode, wmax, part, s);
    case 11: return launch_spread_n<11>(pl, nchunks, w, wmode, wmax, part, s);
    case 12: return launch_spread_n<12>(pl, nchunks, w, wmode, wmax, part, s);
    case 13: return launch_spread_n<13>(pl, nchunks, w, wmode, wmax, part, s);
    default:
      set_error("dsht: the spreading scatter supports Nl <= " + std::to_string(kSpreadMaxNl));
      return DSHT_EINVAL;
  }
}

}  // namespace dsht
