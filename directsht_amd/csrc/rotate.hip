// Polar branch rotation and final combine: alm = a_equ + D(R^-1) a_pol.
//
// The reference rotates the polar points by 90 degrees about y with healpy,
// transforms them, and rotates the result back with healpy's
// Rotator.rotate_alm (sht/sht.py:134-139).  With the points rotated by
// R = R_y(+pi/2) (stage b), the original alm are
//   a_lm = sum_{m'=-l..l} d^l_{m'm}(pi/2) a'_{lm'},  a'_{l,-m'} = (-1)^m' conj(a'_{lm'})
// and, with d^l_{-m',m}(pi/2) = (-1)^{l+m} d^l_{m',m}(pi/2),
//   Re a_lm = d_{0m} Re a'_0 + 2 sum_{m'>0, l+m+m' even} d_{m'm} Re a'_{m'}
//   Im a_lm = d_{0m} Im a'_0 + 2 sum_{m'>0, l+m+m' odd}  d_{m'm} Im a'_{m'}.
// Column m of d^l(pi/2) is generated by the downward three-term recursion
// in m' (Trapani & Navaza, Acta Cryst. A62 (2006) 262)
//   d_{m'-1,m} = [2m d_{m',m} - A(m') d_{m'+1,m}] / B(m'),
//   A = sqrt((l-m')(l+m'+1)),  B = sqrt((l+m')(l-m'+1)),
// from d_{l,m} = (-1)^{l-m} sqrt(C(2l, l+m)) / 2^l, which always runs in the
// direction in which the wanted solution grows.  The column carries a
// power-of-two exponent so starting values below the FP64 range (l > 1022)
// survive.  The starting values and A/B, 1/B depend only on l, so the plan
// precomputes them once.
//
// One workgroup per l, one thread per m; FP64 throughout.  If a polar point
// failed the basic_sht range check after rotation, the reference returns
// zeros for the whole transform (sht/sht.py:141-149); so does this kernel.
#include <mutex>

#include "common.h"
#include "plan.h"

namespace dsht {

constexpr double kScaleHi = 0x1p400;
constexpr double kScaleLo = 0x1p-400;

// G sets per workgroup share the d^l(pi/2) columns generated by the
// recursion; each set has its own a_equ / a_pol ([set][Nlm][2]) and its own
// status flags (flags[set * flag_stride], flag_stride 0 when all sets come
// from the same points).
template <int G>
__global__ __launch_bounds__(256) void k_rotate_combine(const double* __restrict__ a_equ,
                                                        const double* __restrict__ a_pol,
                                                        double* __restrict__ alm, int Nl, int nsets,
                                                        const int* __restrict__ flags,
                                                        int flag_stride, int* __restrict__ status,
                                                        const double* __restrict__ rot_top,
                                                        const int* __restrict__ rot_exp,
                                                        const double* __restrict__ rot_AB,
                                                        const double* __restrict__ rot_iB) {
  const int ell = Nl - 1 - (int)blockIdx.x;  // the longest recursions first
  const int s0 = blockIdx.y * G;
  const int64_t Nlm = (int64_t)Nl * (Nl + 1) / 2;
  // per set: 0 = not in this block, 1 = zeros (polar range failure),
  // 2 = equatorial part only (no polar point), 3 = rotate and combine
  int mode[G];
  bool any_rot = false;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int set = s0 + g;
    mode[g] = 0;
    if (set < nsets) {
      const int fl = flags[set * flag_stride];
      const int npol = flags[set * flag_stride + 1];
      if (status && ell == 0 && threadIdx.x == 0) status[set] = fl;
      mode[g] = (fl & DSHT_STATUS_POLAR_RANGE) ? 1 : (npol == 0 ? 2 : 3);
      any_rot |= mode[g] == 3;
    }
  }
  extern __shared__ double lds[];
  double* sAB = lds;               // A(m')/B(m')
  double* sB = lds + (ell + 1);    // 1/B(m')
  double* sx = lds + 2 * (ell + 1);  // [g][re|im][m']: a'_{l m'} of the sets
  const int64_t tri = (int64_t)ell * (ell + 1) / 2;
  if (any_rot) {
    for (int mp = threadIdx.x; mp <= ell; mp += blockDim.x) {
      sAB[mp] = rot_AB[tri + mp];
      sB[mp] = rot_iB[tri + mp];
      const int64_t idx = hp_index(ell, mp, Nl);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const bool r = mode[g] == 3;
        const double* ap = a_pol + ((int64_t)(s0 + g) * Nlm + idx) * 2;
        sx[(2 * g) * (ell + 1) + mp] = r ? ap[0] : 0.0;
        sx[(2 * g + 1) * (ell + 1) + mp] = r ? ap[1] : 0.0;
      }
    }
    __syncthreads();
  }
  // kChains values of m per thread (m0, m0 + blockDim, ...): independent
  // recursions interleaved step by step, so the dependent FMA of one hides
  // behind the others (a lone recursion is latency-bound)
  constexpr int kChains = 4;
  for (int m0 = threadIdx.x; m0 <= ell; m0 += kChains * blockDim.x) {
    int mm[kChains];
    bool act[kChains];
    double accE[kChains][G], accO[kChains][G], d0[kChains], cur[kChains], nxt[kChains], twom[kChains];
    int E[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      mm[c] = m0 + c * (int)blockDim.x;
      act[c] = mm[c] <= ell;
      d0[c] = 0.0;
#pragma unroll
      for (int g = 0; g < G; ++g) accE[c][g] = accO[c][g] = 0.0;
    }
    if (any_rot) {
      // Step s = l - m' contributes d_{m',m} * Re a'_{m'} when l + m + m' is
      // even and d_{m',m} * Im a'_{m'} when it is odd, i.e. by the parity of
      // m + s: even steps use xe, odd steps xo (no divergent branch).  The
      // chains of a thread have the same parity (blockDim is even).
      const int pe = (m0 & 1), po = 1 - (m0 & 1);  // re|im row of even / odd steps
#pragma unroll
      for (int c = 0; c < kChains; ++c) {
        twom[c] = 2.0 * mm[c];
        cur[c] = act[c] ? rot_top[tri + mm[c]] : 0.0;  // inactive chains stay 0
        E[c] = act[c] ? rot_exp[tri + mm[c]] : 0;
        nxt[c] = 0.0;
      }
      int mp = ell;
      // d_{mp-1} = d_mp * (2m/B) - d_{mp+1} * (A/B): one dependent FMA per step
      while (mp >= 2) {
        const int pairs = min(4, mp >> 1);
        for (int q = 0; q < pairs; ++q) {
          const double be = sB[mp], ae = sAB[mp], bo = sB[mp - 1], ao = sAB[mp - 1];
          double xe[G], xo[G];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            xe[g] = sx[(2 * g + pe) * (ell + 1) + mp];
            xo[g] = sx[(2 * g + po) * (ell + 1) + mp - 1];
          }
#pragma unroll
          for (int c = 0; c < kChains; ++c) {
            double dv = ldexp(cur[c], E[c]);
#pragma unroll
            for (int g = 0; g < G; ++g) accE[c][g] = fma(dv, xe[g], accE[c][g]);
            double nw = fma(cur[c], twom[c] * be, -nxt[c] * ae);
            nxt[c] = cur[c];
            cur[c] = nw;
            dv = ldexp(cur[c], E[c]);
#pragma unroll
            for (int g = 0; g < G; ++g) accO[c][g] = fma(dv, xo[g], accO[c][g]);
            nw = fma(cur[c], twom[c] * bo, -nxt[c] * ao);
            nxt[c] = cur[c];
            cur[c] = nw;
          }
          mp -= 2;
        }
        // at most 8 steps since the last check: growth per step is below 3l,
        // so values stay below 2^400 * (3l)^8 < 2^1023
#pragma unroll
        for (int c = 0; c < kChains; ++c)
          if (fmax(fabs(cur[c]), fabs(nxt[c])) > kScaleHi) {
            cur[c] *= kScaleLo;
            nxt[c] *= kScaleLo;
            E[c] += 400;
          }
      }
#pragma unroll
      for (int c = 0; c < kChains; ++c) {
        if (mp == 1) {
          const double dv = ldexp(cur[c], E[c]);
#pragma unroll
          for (int g = 0; g < G; ++g) accE[c][g] = fma(dv, sx[(2 * g + pe) * (ell + 1) + 1], accE[c][g]);
          cur[c] = fma(cur[c], twom[c] * sB[1], -nxt[c] * sAB[1]);
        }
        d0[c] = ldexp(cur[c], E[c]);  // d_{0,m}
      }
    }
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      if (!act[c]) continue;
      const int m = mm[c];
      const int64_t idx = hp_index(ell, m, Nl);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (mode[g] == 0) continue;
        const int64_t o = ((int64_t)(s0 + g) * Nlm + idx) * 2;
        double re = 0.0, im = 0.0;
        if (mode[g] >= 2) {
          re = a_equ[o];
          im = a_equ[o + 1];
        }
        if (mode[g] == 3) {
          const double sr = (m & 1) ? accO[c][g] : accE[c][g];
          const double si = (m & 1) ? accE[c][g] : accO[c][g];
          re += fma(d0[c], sx[(2 * g) * (ell + 1)], 2.0 * sr);
          im += fma(d0[c], sx[(2 * g + 1) * (ell + 1)], 2.0 * si);
        }
        alm[o] = re;
        alm[o + 1] = im;
      }
    }
  }
}

// Between the healpy order (m-major: idx = m(2Nl-1-m)/2 + l) and the l-major
// order (idx = l(l+1)/2 + m) of nsets arrays [set][Nlm] of complex doubles,
// through 32 x 32 (l, m) tiles in LDS so that both sides are read and written
// as contiguous stretches (in healpy order, consecutive l at fixed m; in
// l-major order, consecutive m at fixed l).  The batched rotation works on
// whole rows l, which are scattered in the healpy order.
template <bool kToLMajor>
__global__ __launch_bounds__(256) void k_tri_transpose(const double2* __restrict__ in,
                                                       double2* __restrict__ out, int Nl) {
  __shared__ double2 tile[32][33];  // [m - m0][l - l0]
  const int l0 = 32 * blockIdx.x, m0 = 32 * blockIdx.y;
  if (m0 > l0 + 31) return;  // tile above the diagonal: no m <= l
  const int64_t Nlm = (int64_t)Nl * (Nl + 1) / 2;
  const double2* src = in + blockIdx.z * Nlm;
  double2* dst = out + blockIdx.z * Nlm;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // read: healpy order walks l at fixed m, l-major order walks m at fixed l
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ty + 8 * r;
    const int l = kToLMajor ? l0 + tx : l0 + i;
    const int m = kToLMajor ? m0 + i : m0 + tx;
    if (l < Nl && m <= l) {
      const int64_t idx = kToLMajor ? hp_index(l, m, Nl) : (int64_t)l * (l + 1) / 2 + m;
      tile[m - m0][l - l0] = src[idx];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = ty + 8 * r;
    const int l = kToLMajor ? l0 + i : l0 + tx;
    const int m = kToLMajor ? m0 + tx : m0 + i;
    if (l < Nl && m <= l) {
      const int64_t idx = kToLMajor ? (int64_t)l * (l + 1) / 2 + m : hp_index(l, m, Nl);
      dst[idx] = tile[m - m0][l - l0];
    }
  }
}

// Many sets: G sets per workgroup (up to 32), one thread per m (rounds of
// 256), the recursion for each (l, m) generated once and applied to all G
// sets.  The a'_{l m'} of the G sets are streamed through LDS in chunks of
// kRotChunk values of m' (all threads walk m' = l .. 1 in the same order), so
// the LDS footprint does not grow with l.
constexpr int kRotChunk = 32;

// Inputs and output in l-major order (k_tri_transpose): a_pol of the sets,
// and a_equ, which is overwritten with the result (each thread reads its own
// (l, m) entries before writing them).
template <int G>
__global__ __launch_bounds__(256) void k_rotate_sets(const double* a_equ,
                                                     const double* __restrict__ a_pol,
                                                     double* alm, int Nl, int nsets,
                                                     const int* __restrict__ flags, int flag_stride,
                                                     int* __restrict__ status,
                                                     const double* __restrict__ rot_top,
                                                     const int* __restrict__ rot_exp,
                                                     const double* __restrict__ rot_AB,
                                                     const double* __restrict__ rot_iB) {
  const int ell = Nl - 1 - (int)blockIdx.x;  // the longest recursions first
  const int s0 = blockIdx.y * G;
  const int64_t Nlm = (int64_t)Nl * (Nl + 1) / 2;
  // a'_{l m'}, m' = hi - row, as [row][re|im][set]: a lane reads its
  // component for two sets per 16-byte LDS read
  __shared__ __attribute__((aligned(16))) double sv[1][kRotChunk][2][G];
  __shared__ double sAB[1][kRotChunk], sBi[1][kRotChunk];
  __shared__ int smode[G];
  if (threadIdx.x < G) {
    const int set = s0 + threadIdx.x;
    int md = 0;  // 0 not in block, 1 zeros, 2 equatorial only, 3 rotate and combine
    if (set < nsets) {
      const int fl = flags[set * flag_stride];
      const int npol = flags[set * flag_stride + 1];
      if (status && ell == 0) status[set] = fl;
      md = (fl & DSHT_STATUS_POLAR_RANGE) ? 1 : (npol == 0 ? 2 : 3);
    }
    smode[threadIdx.x] = md;
  }
  __syncthreads();
  bool any_rot = false;
#pragma unroll
  for (int g = 0; g < G; ++g) any_rot |= smode[g] == 3;
  const int64_t tri = (int64_t)ell * (ell + 1) / 2;
  for (int m0 = 0; m0 <= ell; m0 += blockDim.x) {
    const int m = m0 + threadIdx.x;
    const bool live = m <= ell;
    double accE[G], accO[G];
#pragma unroll
    for (int g = 0; g < G; ++g) accE[g] = accO[g] = 0.0;
    double cur = live ? rot_top[tri + m] : 0.0, nxt = 0.0;
    int E = live ? rot_exp[tri + m] : 0;
    const double twom = 2.0 * m;
    const int cm = m & 1;  // component of even steps: re for even m, im for odd m
    if (any_rot) {
      for (int hi = ell; hi >= 1; hi -= kRotChunk) {
        const int lo = max(1, hi - kRotChunk + 1);
        const int n = hi - lo + 1;
        const int b = 0;
        __syncthreads();  // the previous chunk has been consumed
        // consecutive threads take consecutive m' of one set: a contiguous
        // stretch in l-major order
        for (int e = threadIdx.x; e < n * G; e += blockDim.x) {
          const int row = e % n, g = e / n;
          double2 v = make_double2(0.0, 0.0);
          if (smode[g] == 3)
            v = *reinterpret_cast<const double2*>(a_pol + ((int64_t)(s0 + g) * Nlm + tri + hi - row) * 2);
          sv[b][row][0][g] = v.x;
          sv[b][row][1][g] = v.y;
        }
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
          sAB[b][e] = rot_AB[tri + hi - e];
          sBi[b][e] = rot_iB[tri + hi - e];
        }
        __syncthreads();
        if (live) {
          for (int row = 0; row < n; ++row) {
            // step s = l - m' (m' = hi - row): d_{m',m} times Re a' if s + m is
            // even and Im a' if odd; even steps accumulate in accE, odd in accO
            const double dv = ldexp(cur, E);
            typedef double d2 __attribute__((ext_vector_type(2)));
            if (((ell - hi + row) & 1) == 0) {  // uniform branch
              const d2* v = reinterpret_cast<const d2*>(&sv[b][row][cm][0]);
#pragma unroll
              for (int g = 0; g < G; g += 2) {
                const d2 a = v[g / 2];
                accE[g] = fma(dv, a.x, accE[g]);
                accE[g + 1] = fma(dv, a.y, accE[g + 1]);
              }
            } else {
              const d2* v = reinterpret_cast<const d2*>(&sv[b][row][cm ^ 1][0]);
#pragma unroll
              for (int g = 0; g < G; g += 2) {
                const d2 a = v[g / 2];
                accO[g] = fma(dv, a.x, accO[g]);
                accO[g + 1] = fma(dv, a.y, accO[g + 1]);
              }
            }
            // d_{m'-1} = d_m' * (2m/B) - d_{m'+1} * (A/B); values stay below 2^1023
            // since they are rescaled as soon as they pass 2^400
            const double nw = fma(cur, twom * sBi[b][row], -nxt * sAB[b][row]);
            nxt = cur;
            cur = nw;
            if (fmax(fabs(cur), fabs(nxt)) > kScaleHi) {
              cur *= kScaleLo;
              nxt *= kScaleLo;
              E += 400;
            }
          }
        }
      }
    }
    if (live) {
      const double d0 = ldexp(cur, E);  // d_{0,m}
      const int64_t idx = tri + m, idx0 = tri;  // l-major
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int md = smode[g];
        if (md == 0) continue;
        const int64_t o = ((int64_t)(s0 + g) * Nlm + idx) * 2;
        double re = 0.0, im = 0.0;
        if (md >= 2) {
          re = a_equ[o];
          im = a_equ[o + 1];
        }
        if (md == 3) {
          const double* a0 = a_pol + ((int64_t)(s0 + g) * Nlm + idx0) * 2;
          const double sr = cm ? accO[g] : accE[g];
          const double si = cm ? accE[g] : accO[g];
          re += fma(d0, a0[0], 2.0 * sr);
          im += fma(d0, a0[1], 2.0 * si);
        }
        alm[o] = re;
        alm[o + 1] = im;
      }
    }
  }
}

// a_pol -> a_tmp and a_equ -> a_pol in l-major order; the rotation and
// combine in place in a_pol; back to the healpy order into alm.
template <int G>
static int launch_rotate_sets(Plan& pl, int nsets, const int* flags, int flag_stride, double* alm,
                              int* status, hipStream_t s) {
  Workspace& ws = pl.ws;
  if (!ws.a_tmp || nsets > ws.cap_sets) {
    set_error("dsht: batched rotation without its workspace");
    return DSHT_EINVAL;
  }
  const int nt = (pl.Nl + 31) / 32;
  const dim3 tg(nt, nt, nsets);
  k_tri_transpose<true><<<tg, 256, 0, s>>>((const double2*)ws.a_pol, (double2*)ws.a_tmp, pl.Nl);
  DSHT_LAUNCH_CHECK();
  k_tri_transpose<true><<<tg, 256, 0, s>>>((const double2*)ws.a_equ, (double2*)ws.a_pol, pl.Nl);
  DSHT_LAUNCH_CHECK();
  dim3 grid(pl.Nl, (nsets + G - 1) / G);
  k_rotate_sets<G><<<grid, 256, 0, s>>>(ws.a_pol, ws.a_tmp, ws.a_pol, pl.Nl, nsets, flags,
                                        flag_stride, status, pl.rot_top, pl.rot_exp, pl.rot_AB,
                                        pl.rot_iB);
  DSHT_LAUNCH_CHECK();
  k_tri_transpose<false><<<tg, 256, 0, s>>>((const double2*)ws.a_pol, (double2*)alm, pl.Nl);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

// alm[set] = a_equ[set] + D(R^-1) a_pol[set] for nsets sets; status[set]
// (may be NULL) receives the set's DSHT_STATUS_* bits.
int launch_rotate_combine(Plan& pl, int nsets, const int* flags, int flag_stride, double* alm,
                          int* status, hipStream_t s) {
  if (nsets == 1) {
    const size_t lds = sizeof(double) * 4 * (size_t)pl.Nl;
    if (lds > 64 * 1024) {  // beyond the default dynamic LDS of a launch
      static std::once_flag once;
      static hipError_t attr = hipSuccess;
      std::call_once(once, [] {
        attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rotate_combine<1>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double) * 4 * kMaxNl));
      });
      DSHT_HIP(attr);
    }
    k_rotate_combine<1><<<pl.Nl, 256, lds, s>>>(pl.ws.a_equ, pl.ws.a_pol, alm, pl.Nl, 1, flags,
                                                flag_stride, status, pl.rot_top, pl.rot_exp,
                                                pl.rot_AB, pl.rot_iB);
    DSHT_LAUNCH_CHECK();
    return DSHT_OK;
  }
  // at most 16 sets per workgroup: 32 would need two waves per SIMD (FP64
  // accumulators) and runs slower than recomputing the d-columns per 16 sets
  if (nsets > 8) return launch_rotate_sets<16>(pl, nsets, flags, flag_stride, alm, status, s);
  if (nsets > 4) return launch_rotate_sets<8>(pl, nsets, flags, flag_stride, alm, status, s);
  return launch_rotate_sets<4>(pl, nsets, flags, flag_stride, alm, status, s);
}

}  // namespace dsht
