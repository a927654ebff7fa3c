// Angular power spectra of alm (healpy order) on the device: the reference's
// DirectSHT.alm2cl (sht/sht.py:95-109),
//   C_l = [ (Re a_l0)^2 + 2 sum_{m>0} |a_lm|^2 ] / (2l + 1),
// and, for pairs of sets, the cross-spectra its callers take with healpy's
// alm2cl(alm1, alm2) (notebooks/analyzing_mocks.ipynb), with the same m = 0
// convention (real parts only):
//   C^{st}_l = [ Re a^s_l0 Re a^t_l0 + 2 sum_{m>0} Re(a^s_lm conj(a^t_lm)) ] / (2l + 1).
// C^{ss} is the auto-spectrum of set s.
//
// Cross-spectra: one workgroup per l (the longest rows first, four
// consecutive l's per XCD).  The row's a_lm
// of all sets are staged through LDS kSpecM orders at a time ([m][set]); each
// thread accumulates its pairs (s, t) in FP64 in order of m.  Auto-spectra
// alone take a coalesced kernel (k_alm2cl_auto).
#include "common.h"
#include "plan.h"

namespace dsht {

constexpr int kSpecM = 32;                                      // orders per LDS chunk
constexpr int kSpecPairs = (kMaxSets * (kMaxSets + 1) / 2 + 255) / 256;  // pairs per thread

__global__ __launch_bounds__(256) void k_alm2cl(const double2* __restrict__ alm, int64_t ld,
                                                int Nl, int nsets, int cross,
                                                double* __restrict__ cl) {
  __shared__ double2 sa[kSpecM][kMaxSets];
  // Four consecutive l's on one XCD (block b runs on XCD b % 8): their a_lm
  // for a fixed m are adjacent 16-byte values of one healpy row, so each line
  // is read from HBM once and shared through that XCD's L2.  Groups of four
  // are interleaved over the XCDs, longest rows first.
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int ell = Nl - 1 - (((slot >> 2) * 8 + xcd) * 4 + (slot & 3));
  if (ell < 0) return;
  const int npairs = cross ? nsets * (nsets + 1) / 2 : nsets;
  // this thread's pairs p = tid + 256 q: (s, t), s <= t, in row order
  int ps[kSpecPairs], pt[kSpecPairs];
#pragma unroll
  for (int q = 0; q < kSpecPairs; ++q) {
    int p = threadIdx.x + 256 * q;
    ps[q] = pt[q] = -1;
    if (p >= npairs) continue;
    if (!cross) {
      ps[q] = pt[q] = p;
      continue;
    }
    int s = 0;
    while (p >= nsets - s) {  // row s holds pairs (s, s .. nsets-1)
      p -= nsets - s;
      ++s;
    }
    ps[q] = s;
    pt[q] = s + p;
  }
  double acc[kSpecPairs];
#pragma unroll
  for (int q = 0; q < kSpecPairs; ++q) acc[q] = 0.0;
  for (int m0 = 0; m0 <= ell; m0 += kSpecM) {
    const int nm = min(kSpecM, ell + 1 - m0);
    __syncthreads();  // the previous chunk has been used
    for (int e = threadIdx.x; e < nm * nsets; e += blockDim.x) {
      const int mi = e % nm, s = e / nm;
      sa[mi][s] = alm[(int64_t)s * ld + hp_index(ell, m0 + mi, Nl)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kSpecPairs; ++q) {
      if (ps[q] < 0) continue;
      double a = acc[q];
      for (int mi = 0; mi < nm; ++mi) {
        const double2 x = sa[mi][ps[q]], y = sa[mi][pt[q]];
        if (m0 + mi == 0)
          a = fma(x.x, y.x, a);
        else
          a = fma(2.0 * x.x, y.x, fma(2.0 * x.y, y.y, a));
      }
      acc[q] = a;
    }
  }
  const double inv = 1.0 / (2.0 * ell + 1.0);
#pragma unroll
  for (int q = 0; q < kSpecPairs; ++q)
    if (ps[q] >= 0) cl[(int64_t)(threadIdx.x + 256 * q) * Nl + ell] = acc[q] * inv;
}

// Auto-spectra: block (64 consecutive l, one set); lane = l, the four
// wavefronts take m = w, w + 4, ...  For fixed m the l's of a healpy row are
// contiguous, so each wavefront load is one coalesced 1 KB read and every
// alm is read once.  The four partial sums are added in wavefront order.
__global__ __launch_bounds__(256) void k_alm2cl_auto(const double2* __restrict__ alm, int64_t ld,
                                                     int Nl, double* __restrict__ cl) {
  __shared__ double part[4][64];
  const int nlc = (Nl + 63) / 64;
  const int lc = nlc - 1 - (int)blockIdx.x;  // longest rows first
  const int s = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ell = lc * 64 + lane;
  const int mend = min(lc * 64 + 63, Nl - 1);
  const double2* a = alm + (int64_t)s * ld;
  double acc = 0.0;
  const int mtop = ell < Nl ? min(ell, mend) : -1;  // this lane's last m
  for (int m0 = w; m0 <= mend; m0 += 32) {           // eight loads in flight
    double2 x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int m = m0 + 4 * u;
      x[u] = m <= mtop ? a[hp_index(ell, m, Nl)] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int m = m0 + 4 * u;
      acc = m == 0 ? fma(x[u].x, x[u].x, acc)
                   : fma(2.0 * x[u].x, x[u].x, fma(2.0 * x[u].y, x[u].y, acc));
    }
  }
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0 && ell < Nl) {
    const double t = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    cl[(int64_t)s * Nl + ell] = t / (2.0 * ell + 1.0);
  }
}

int launch_alm2cl(int Nl, int nsets, const double* alm, int64_t ld, double* cl, int cross,
                  hipStream_t s) {
  if (!cross) {
    k_alm2cl_auto<<<dim3((Nl + 63) / 64, nsets), 256, 0, s>>>(
        reinterpret_cast<const double2*>(alm), ld, Nl, cl);
    DSHT_LAUNCH_CHECK();
    return DSHT_OK;
  }
  k_alm2cl<<<(Nl + 31) / 32 * 32, 256, 0, s>>>(reinterpret_cast<const double2*>(alm), ld, Nl, nsets,
                                               cross, cl);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
