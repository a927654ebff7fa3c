// Wigner 3j symbols (l1 l2 l3; 0 0 0) and the pseudo-C_l mode-coupling matrix
// of a window function, on the device.
//
// Both use the closed form of the m = 0 symbol (Racah): with J = l1 + l2 + l3
// even, g = J/2 and
//     G(p) = (2p)! / (4^p (p!)^2) = prod_{q=1..p} (2q - 1)/(2q),
//     H(g) = 1 / (G(g) (2g + 1)),
//     (l1 l2 l3; 0 0 0)^2 = G(g - l1) G(g - l2) G(g - l3) H(g),
// sign (-1)^g, zero for odd J or off the triangle |l1 - l2| <= l3 <= l1 + l2.
// G and H are computed once per device on the host in extended precision
// (80-bit x87 on the x86 host) and kept in HBM (96 KB), so a symbol costs
// three products and a square root instead of the reference's recursion of up
// to j3 steps (sht/threej000.py:9-42, csht/threej000_helper.c:28-56).
//
// Mode coupling (MaskDeconvolution.get_M, sht/mask_deconvolution.py:171-192):
//     M[l1][l2] = (2 l2 + 1)/(4 pi) sum_{l3} (2 l3 + 1) W(l3) (l1 l2 l3; 0 0 0)^2.
// The sum S(l1, l2) is symmetric, so only l2 <= l1 is summed; with
// d = l1 - l2 and l3 = d + 2k (k = 0 .. l2), g = l1 + k and
//     (l1 l2 l3; 0 0 0)^2 = G(k) G(d + k) G(l2 - k) H(l1 + k).
// A block takes 64 consecutive l2 (one per lane) and 4 or 8 rows l1, its four
// wavefronts a quarter of the k's each; G(k) G(l2 - k) is shared by the rows,
// H(l1 + k) is uniform per row.  Each wavefront sums its terms in
// ascending l3, as the reference does.
#include <mutex>
#include <vector>

#include "common.h"
#include "plan.h"

namespace dsht {

constexpr int kTjMaxN = 4096;      // largest table size N and mode-coupling Nl
constexpr int kTjG = 4096;         // G(p), p < kTjG
constexpr int kTjH = 8192;         // H(g), g < kTjH  (g <= 3(N-1)/2 and <= 2Nl-2)
constexpr int kTjChunk = 4096;     // table entries per work chunk (16 per thread)
constexpr int kMcRowsMax = 8;      // rows l1 per mode-coupling tile: 4, or 8 from kMcRows8Nl
constexpr int kMcRows8Nl = 1536;   // (more independent sums per thread, fewer blocks)
constexpr int kMcCols = 64;        // columns l2 per tile (one per lane)

namespace {
std::mutex g_tj_mu;
std::vector<double*> g_tj_coef;    // per device: G[kTjG] then H[kTjH]
}  // namespace

// G and H for this device (built on its first use; never freed).
static int tj_coefficients(const double** out) {
  int dev = 0;
  DSHT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_tj_mu);
  if ((int)g_tj_coef.size() <= dev) g_tj_coef.resize(dev + 1, nullptr);
  if (!g_tj_coef[dev]) {
    std::vector<double> h(kTjG + kTjH);
    long double G = 1.0L;
    for (int p = 0; p < kTjH; ++p) {
      if (p > 0) G *= (long double)(2 * p - 1) / (long double)(2 * p);
      if (p < kTjG) h[p] = (double)G;
      h[kTjG + p] = (double)(1.0L / (G * (long double)(2 * p + 1)));
    }
    double* d = nullptr;
    DSHT_HIP(hipMalloc(&d, h.size() * sizeof(double)));
    hipError_t e = hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      set_error(std::string("hipMemcpy of the 3j coefficients: ") + hipGetErrorString(e));
      return DSHT_EHIP;
    }
    g_tj_coef[dev] = d;
  }
  *out = g_tj_coef[dev];
  return DSHT_OK;
}

__host__ __device__ inline int64_t tetra(int64_t j) { return j * (j + 1) * (j + 2) / 6; }

// store[tetra(j1) + j2(j2+1)/2 + j3] = (j1 j2 j3; 0 0 0), N > j1 >= j2 >= j3.
// The flat index space is cut into chunks of kTjChunk entries; a block walks
// chunks with a grid stride, each thread 16 entries 256 apart (coalesced
// stores): the first entry's (j1, j2, j3) from the chunk's j1 and a square
// root, the next ones by stepping j3 by 256 with carries into j2 and j1.
__global__ __launch_bounds__(256) void k_threej000_table(int N, int64_t total,
                                                         const double* __restrict__ G,
                                                         const double* __restrict__ H,
                                                         double* __restrict__ store) {
  for (int64_t c0 = (int64_t)blockIdx.x * kTjChunk; c0 < total;
       c0 += (int64_t)gridDim.x * kTjChunk) {
    // j1 of the chunk's first entry: the largest j with tetra(j) <= c0; the
    // row j1 spans [rs, re) of the flat index
    int64_t j1 = (int64_t)cbrt(6.0 * (double)c0);
    while (j1 > 0 && tetra(j1) > c0) --j1;
    while (tetra(j1 + 1) <= c0) ++j1;
    int64_t rs = tetra(j1), re = tetra(j1 + 1);
    // this thread's first entry: (j1, j2, j3) from its offset in row j1; the
    // next entries, 256 further on, by stepping j3 and carrying into j2, j1
    int64_t i = c0 + threadIdx.x;
    if (i >= total) continue;
    while (i >= re) {
      ++j1;
      rs = re;
      re += (j1 + 1) * (j1 + 2) / 2;
    }
    int t0 = (int)(i - rs);  // < N(N+1)/2 < 2^24
    int j2 = (int)((sqrtf(8.0f * (float)t0 + 1.0f) - 1.0f) * 0.5f);
    while (j2 > 0 && j2 * (j2 + 1) / 2 > t0) --j2;
    while ((j2 + 1) * (j2 + 2) / 2 <= t0) ++j2;
    int j3 = t0 - j2 * (j2 + 1) / 2;
    for (int r = 0; r < kTjChunk / 256; ++r, i += 256) {
      if (r > 0) {
        if (i >= total) break;
        j3 += 256;
        while (j3 > j2) {  // carry: row j2 holds j3 = 0 .. j2
          j3 -= j2 + 1;
          if (++j2 > j1) {  // triangle j1 holds j2 = 0 .. j1
            j2 = 0;
            ++j1;
          }
        }
      }
      const int J = (int)j1 + j2 + j3;
      double v = 0.0;
      if (!(J & 1) && j3 >= (int)j1 - j2) {
        const int g = J >> 1;
        v = sqrt(G[g - (int)j1] * G[g - j2] * G[g - j3] * H[g]);
        if (g & 1) v = -v;
      }
      store[i] = v;
    }
  }
}

// One tile: rows l1 in [R0, R0 + kMcRows), columns l2 in [C0, C0 + 64); entries
// with l2 <= l1 are summed and written to M[l1][l2] and M[l2][l1].  The four
// wavefronts of a block split the sum (k = w, w + 4, ...) and their partial
// sums are added in wavefront order.  Tiles are taken with the longest sums
// (largest C0) first.
// kLds (Nl <= kMcLdsNl): the block's range of G and its window of
// (2 l3 + 1) W(l3) are staged in LDS first, and the per-row gathers read LDS
// instead of the L1 cache.
constexpr int kMcLdsNl = 2048;

template <bool kLds, int kMcRows>
__global__ __launch_bounds__(256) void k_mode_coupling(int Nl, const double* __restrict__ W,
                                                       int64_t nW,
                                                       const double* __restrict__ G,
                                                       const double* __restrict__ H,
                                                       double* __restrict__ M) {
  __shared__ double part[4][kMcRows][kMcCols];
  __shared__ double sG[kLds ? kMcLdsNl : 1];
  __shared__ double sW[kLds ? 2 * kMcLdsNl + 2 * kMcCols + 2 * kMcRowsMax : 1];
  const int TC = (Nl + kMcCols - 1) / kMcCols, TR = (Nl + kMcRows - 1) / kMcRows;
  // block -> (tc, tr): tc from TC-1 down, tr from the first row tile that
  // reaches the diagonal (R0 + kMcRows - 1 >= C0) to TR-1
  int b = blockIdx.x, tc = TC - 1, tr0 = 0;
  for (;; --tc) {
    tr0 = (tc * kMcCols) / kMcRows;
    const int n = TR - tr0;
    if (b < n) break;
    b -= n;
  }
  const int C0 = tc * kMcCols, R0 = (tr0 + b) * kMcRows;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar k loads
  const int l2 = C0 + lane;
  const int nl3 = (int)min<int64_t>(nW, 2 * (int64_t)Nl - 1);  // W(l3) = 0 beyond
  int l1[kMcRows], d[kMcRows];
  bool on[kMcRows];
  double acc[kMcRows];
#pragma unroll
  for (int r = 0; r < kMcRows; ++r) {
    l1[r] = R0 + r;
    d[r] = l1[r] - l2;
    on[r] = l1[r] < Nl && l2 < Nl && d[r] >= 0;
    acc[r] = 0.0;
  }
  // largest l2 summed here; nothing to sum for an empty window
  const int kend = nl3 > 0 ? min(C0 + kMcCols - 1, R0 + kMcRows - 1) : -1;
  // window of l3 = l1 - l2 + 2k that this tile can reach: [wlo, whi)
  const int wlo = max(0, R0 - C0 - (kMcCols - 1));
  const int whi = min(nl3, R0 + C0 + kMcRows + kMcCols);
  if constexpr (kLds) {
    const int gcount = min(Nl, max(C0 + kMcCols, R0 + kMcRows));  // G indices used
    for (int i = threadIdx.x; i < gcount; i += blockDim.x) sG[i] = G[i];
    for (int i = threadIdx.x; i < whi - wlo; i += blockDim.x)
      sW[i] = (double)(2 * (wlo + i) + 1) * W[wlo + i];
    __syncthreads();
  }
  // Branch-free body: masked terms load from clamped in-bounds indices and add
  // zero, so the loads of all rows are in flight together.
  for (int k = wv; k <= kend; k += 4) {
    const bool kin = k <= l2 && l2 < Nl;
    const double q = kLds ? sG[k] * sG[kin ? l2 - k : 0] : G[k] * G[kin ? l2 - k : 0];
#pragma unroll
    for (int r = 0; r < kMcRows; ++r) {
      const int l3 = d[r] + 2 * k;
      const bool v = on[r] && kin && l3 < nl3;
      const double w = kLds ? sW[v ? l3 - wlo : 0] : (double)(2 * l3 + 1) * W[v ? l3 : 0];
      const double h = H[l1[r] < Nl ? l1[r] + k : 0];  // uniform: a scalar load
      const double g = kLds ? sG[v ? d[r] + k : 0] : G[v ? d[r] + k : 0];
      const double tj2 = q * g * h;
      acc[r] += v ? w * tj2 : 0.0;
    }
  }
#pragma unroll
  for (int r = 0; r < kMcRows; ++r) part[wv][r][lane] = acc[r];
  __syncthreads();
  const double four_pi = 4.0 * M_PI;
#pragma unroll
  for (int r = 0; r < kMcRows; ++r) {  // wavefront w writes rows R0 + w, R0 + w + 4, ...
    if ((r & 3) != wv || !on[r]) continue;
    const double S = ((part[0][r][lane] + part[1][r][lane]) + part[2][r][lane]) + part[3][r][lane];
    M[(int64_t)l1[r] * Nl + l2] = (S * (double)(2 * l2 + 1)) / four_pi;
    if (d[r] > 0) M[(int64_t)l2 * Nl + l1[r]] = (S * (double)(2 * l1[r] + 1)) / four_pi;
  }
}

int launch_threej000_table(int N, double* store, hipStream_t s) {
  const double* coef = nullptr;
  int rc = tj_coefficients(&coef);
  if (rc != DSHT_OK) return rc;
  const int64_t total = tetra(N);
  const int64_t chunks = (total + kTjChunk - 1) / kTjChunk;
  const int nb = (int)std::min<int64_t>(chunks, 8 * 1024);
  k_threej000_table<<<nb, 256, 0, s>>>(N, total, coef, coef + kTjG, store);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

template <int kRows>
static int launch_mode_coupling_rows(int Nl, const double* W, int64_t nW, const double* coef,
                                     double* M, hipStream_t s) {
  const int TC = (Nl + kMcCols - 1) / kMcCols, TR = (Nl + kRows - 1) / kRows;
  int nb = 0;
  for (int tc = 0; tc < TC; ++tc) nb += TR - (tc * kMcCols) / kRows;
  if (Nl <= kMcLdsNl)
    k_mode_coupling<true, kRows><<<nb, 256, 0, s>>>(Nl, W, nW, coef, coef + kTjG, M);
  else
    k_mode_coupling<false, kRows><<<nb, 256, 0, s>>>(Nl, W, nW, coef, coef + kTjG, M);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_mode_coupling(int Nl, const double* W, int64_t nW, double* M, hipStream_t s) {
  const double* coef = nullptr;
  int rc = tj_coefficients(&coef);
  if (rc != DSHT_OK) return rc;
  return Nl >= kMcRows8Nl ? launch_mode_coupling_rows<8>(Nl, W, nW, coef, M, s)
                          : launch_mode_coupling_rows<4>(Nl, W, nW, coef, M, s);
}

int tj_max_n() { return kTjMaxN; }

}  // namespace dsht
