// Angular power spectra of alm (healpy order) on the device: the reference's
// DirectSHT.alm2cl (sht/sht.py:95-109),
//   C_l = [ (Re a_l0)^2 + 2 sum_{m>0} |a_lm|^2 ] / (2l + 1),
// and, for pairs of sets, the cross-spectra its callers take with healpy's
// alm2cl(alm1, alm2) (notebooks/analyzing_mocks.ipynb), with the same m = 0
// convention (real parts only):
//   C^{st}_l = [ Re a^s_l0 Re a^t_l0 + 2 sum_{m>0} Re(a^s_lm conj(a^t_lm)) ] / (2l + 1).
// C^{ss} is the auto-spectrum of set s.
//
// Auto-spectra take a coalesced kernel (k_alm2cl_auto), cross-spectra the
// FP64 matrix cores (k_alm2cl_gram).
#include "common.h"
#include "plan.h"

namespace dsht {

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

// Cross-spectra as Gram matrices on the FP64 matrix cores.  For each l,
//   G_l[s][t] = sum_k X_l[s][k] w_k X_l[t][k],  k = (m, Re|Im),
// w = 1 for Re a_l0, 0 for Im a_l0 and 2 for m > 0 (exact), so that
// C^{st}_l = G_l[s][t] / (2l+1).  The sets form tiles of 16; a workgroup takes
// kGL = 8 consecutive l (one 128-byte healpy line per set and order) and one
// tile pair (i <= j), and wavefront w the degree l0 + w: per step of two
// orders (K = 4 values: Re, Im of m and m + 1) one v_mfma_f64_16x16x4f64 with
// A = X[16 i + lane % 16][k] and B = w_k X[16 j + lane % 16][k],
// k = lane / 16.  The a_lm are staged through LDS kGM orders at a time, each
// chunk loaded into registers two chunks ahead: a wavefront's load
// instruction covers eight whole 128-byte lines (8 l of 8 sets), and the LDS
// layout [tile][m][Re|Im][set][l ^ (set / 2)] keeps both the stores and the
// operand reads (16 sets at one l) free of bank conflicts.  The tile pairs of
// one l-block run on one XCD, so the lines each of them reads come from HBM
// once.  Every sum is formed in a fixed order (no atomics).
constexpr int kGL = 8;   // degrees per workgroup (one wavefront each)
constexpr int kGM = 16;  // orders per staged chunk
typedef double d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(512) void k_alm2cl_gram(const double2* __restrict__ alm, int64_t ld,
                                                     int Nl, int nsets, int nlb, int npt,
                                                     double* __restrict__ cl) {
  __shared__ double sX[2][kGM][2][16][kGL];  // [tile][m][Re|Im][set][l ^ (set / 2)]
  // blocks in groups of 8 l-blocks x npt tile pairs: the tile pairs of an
  // l-block on one XCD (block b runs on XCD b % 8); longest rows first
  const int L = blockIdx.x, per = 8 * npt;
  const int g = L / per, r = L - g * per;
  const int lbr = 8 * g + (r & 7), pt = r >> 3;
  if (lbr >= nlb) return;
  const int lb = nlb - 1 - lbr;
  int ti = 0, tj = pt;  // tile pair number pt -> (ti, tj), ti <= tj
  {
    const int nt = (nsets + 15) / 16;
    while (tj >= nt - ti) {
      tj -= nt - ti;
      ++ti;
    }
    tj += ti;
  }
  const int l0 = kGL * lb;
  const int lmax = min(l0 + kGL - 1, Nl - 1);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int ell = l0 + w;
  const int nq = ti == tj ? 1 : 2;  // tiles to stage
  // staging: element e = tid + 512 v, v < 8: l e % 8, set (e / 8) % 16,
  // m (e / 128) % 16, tile e / 2048
  double2 bufA[8], bufB[8];
  auto load = [&](double2 (&pre)[8], int m0) {
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int e = tid + 512 * v;
      const int li = e & 7, si = (e >> 3) & 15, mi = (e >> 7) & 15, q = e >> 11;
      const int set = 16 * (q ? tj : ti) + si, l = l0 + li, m = m0 + mi;
      pre[v] = make_double2(0.0, 0.0);
      if (q < nq && set < nsets && l <= lmax && m <= l)
        pre[v] = alm[(int64_t)set * ld + hp_index(l, m, Nl)];
    }
  };
  auto stage = [&](const double2 (&pre)[8]) {
    __syncthreads();  // the previous chunk has been used
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int e = tid + 512 * v;
      const int li = e & 7, s2 = (e >> 3) & 15, mi = (e >> 7) & 15, q = e >> 11;
      sX[q][mi][0][s2][li ^ (s2 >> 1)] = pre[v].x;
      sX[q][mi][1][s2][li ^ (s2 >> 1)] = pre[v].y;
    }
    __syncthreads();
  };
  d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
  const int si = lane & 15, kk = lane >> 4;  // operand element (set, k)
  const int qb = nq - 1;                     // the B tile's slot
  const int c = kk & 1, sw = w ^ (si >> 1);
  // all eight steps of two orders (m0 + 2 st, + 1): the operands are read
  // first, then multiplied (entries past this wavefront's l are staged zeros)
  auto compute = [&](int m0) {
    if (ell > lmax || m0 > ell) return;  // wave-uniform
    double a[kGM / 2], b[kGM / 2];
#pragma unroll
    for (int st = 0; st < kGM / 2; ++st) {
      const int mi = 2 * st + (kk >> 1), m = m0 + mi;
      const double wk = m == 0 ? (c == 0 ? 1.0 : 0.0) : 2.0;
      a[st] = sX[0][mi][c][si][sw];
      b[st] = wk * sX[qb][mi][c][si][sw];
    }
#pragma unroll
    for (int st = 0; st < kGM / 2; st += 2) {
      acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[st], b[st], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[st + 1], b[st + 1], acc1, 0, 0, 0);
    }
  };
  // chunks alternate between two register buffers, each loaded two chunks
  // ahead of its use
  load(bufA, 0);
  if (kGM <= lmax) load(bufB, kGM);
  for (int m0 = 0; m0 <= lmax; m0 += 2 * kGM) {
    stage(bufA);
    if (m0 + 2 * kGM <= lmax) load(bufA, m0 + 2 * kGM);
    compute(m0);
    if (m0 + kGM > lmax) break;
    stage(bufB);
    if (m0 + 3 * kGM <= lmax) load(bufB, m0 + 3 * kGM);
    compute(m0 + kGM);
  }
  // D layout (checked on gfx950): lane, register r hold
  // G[16 ti + 4 r + lane / 16][16 tj + lane % 16]
  if (ell <= lmax) {
    const double inv = 1.0 / (2.0 * ell + 1.0);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int s1 = 16 * ti + 4 * rr + kk, t1 = 16 * tj + si;
      if (s1 < nsets && t1 < nsets && s1 <= t1) {
        const int64_t p = (int64_t)s1 * nsets - (int64_t)s1 * (s1 - 1) / 2 + (t1 - s1);
        cl[p * Nl + ell] = (acc0[rr] + acc1[rr]) * inv;
      }
    }
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
  const int nt = (nsets + 15) / 16, npt = nt * (nt + 1) / 2;
  const int nlb = (Nl + kGL - 1) / kGL;
  k_alm2cl_gram<<<(unsigned)((nlb + 7) / 8 * 8 * npt), 512, 0, s>>>(
      reinterpret_cast<const double2*>(alm), ld, Nl, nsets, nlb, npt, cl);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
