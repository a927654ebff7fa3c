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

// healpy order [set][Nlm] -> grouped l-major order [set / G][Nlm][re|im][G]
// (the batched rotation's input: for each (l, m') and component, the G sets
// of a group are 8 G contiguous bytes).  One workgroup per tile of 16 l x 8 m
// of one group: the G sets' tiles are read (consecutive l at fixed m) into
// LDS and written out (consecutive sets, components and m at fixed l) as
// whole 128-byte lines.  Sets past nsets are left unwritten.
template <int G>
__global__ __launch_bounds__(256) void k_tri_group(const double2* __restrict__ in,
                                                   double* __restrict__ out, int Nl, int nsets) {
  __shared__ double2 tile[G][8][16];  // [set][m - m0][l - l0]
  const int l0 = 16 * blockIdx.x, m0 = 8 * blockIdx.y, grp = blockIdx.z;
  if (m0 > l0 + 15) return;  // tile above the diagonal: no m <= l
  const int64_t Nlm = (int64_t)Nl * (Nl + 1) / 2;
  const int t = threadIdx.x;
#pragma unroll
  for (int r = 0; r < G * 128 / 256; ++r) {
    const int e = t + 256 * r;
    const int li = e & 15, mi = (e >> 4) & 7, g = e >> 7;
    const int l = l0 + li, m = m0 + mi, set = grp * G + g;
    if (l < Nl && m <= l && set < nsets) tile[g][mi][li] = in[set * Nlm + hp_index(l, m, Nl)];
  }
  __syncthreads();
  double* dst = out + (int64_t)grp * Nlm * 2 * G;
#pragma unroll
  for (int r = 0; r < G * 256 / 256; ++r) {
    const int e = t + 256 * r;
    const int g = e % G, comp = (e / G) & 1, q = e / (2 * G);
    const int mi = q & 7, li = q >> 3;
    const int l = l0 + li, m = m0 + mi, set = grp * G + g;
    if (l < Nl && m <= l && set < nsets) {
      const double2 v = tile[g][mi][li];
      dst[(((int64_t)l * (l + 1) / 2 + m) * 2 + comp) * G + g] = comp ? v.y : v.x;
    }
  }
}

// Many sets: G sets per workgroup, the recursion for each (l, m) generated
// once and applied to all G sets.  Every lane of a wavefront needs the same
// a'_{l m'} at each step, so the sets' values are read with scalar loads
// (through the constant address space: uniform addresses, read-only data)
// and enter the FMAs as scalar operands; no LDS.  The component a step reads
// (Re or Im) follows the parity of m, so each wavefront takes m of one
// parity: wavefront w of the block holds m = m0 + 2 (64 (w >> 1) + lane) +
// (w & 1), and each thread C such values m, m + 256, ... (chains sharing
// every scalar operand).  Accumulators are kept in the chain's scale 2^E
// (the same power of two as its d values) and rescaled with them, which
// is exact; the true sums are formed at the end.
typedef const double __attribute__((address_space(4)))* kdptr;
typedef const int __attribute__((address_space(4)))* kiptr;

// Inputs: a_pol of the sets in grouped l-major order (k_tri_group), a_equ in
// l-major order (k_tri_transpose), which is overwritten with the result
// (each thread reads its own (l, m) entries before writing them).
template <int G, int C>
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
  const int64_t tri = (int64_t)ell * (ell + 1) / 2;
  const kiptr fl = (kiptr)flags;
  // per set: 0 = not in this block, 1 = zeros (polar range failure),
  // 2 = equatorial part only (no polar point), 3 = rotate and combine
  int mode[G];
  bool any_rot = false;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int set = s0 + g;
    mode[g] = 0;
    if (set < nsets) {
      const int f = fl[(int64_t)set * flag_stride];
      const int npol = fl[(int64_t)set * flag_stride + 1];
      mode[g] = (f & DSHT_STATUS_POLAR_RANGE) ? 1 : (npol == 0 ? 2 : 3);
      any_rot |= mode[g] == 3;
    }
  }
  if (status && ell == 0 && (int)threadIdx.x < G && s0 + (int)threadIdx.x < nsets)
    status[s0 + threadIdx.x] = flags[(int64_t)(s0 + threadIdx.x) * flag_stride];
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int p = w & 1;                     // parity of this wavefront's m
  const int wfirst = 128 * (w >> 1) + p;   // its smallest m - m0
  const int mw = wfirst + 2 * lane;        // this thread's m - m0 (chain 0)
  // a'_{l m'} of set s0 + g: xs[(2 m' + component) G + g] (sets past nsets:
  // unused values)
  const kdptr xs = (kdptr)a_pol + ((int64_t)blockIdx.y * Nlm + tri) * 2 * G;
  const kdptr ab = (kdptr)rot_AB + tri;
  const kdptr ib = (kdptr)rot_iB + tri;
  for (int m0 = 0; m0 + wfirst <= ell; m0 += 256 * C) {
    int mm[C], E[C];
    bool act[C];
    double cur[C], nxt[C], twom[C], accE[C][G], accO[C][G];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      mm[c] = m0 + mw + 256 * c;
      act[c] = mm[c] <= ell;
      cur[c] = act[c] ? rot_top[tri + mm[c]] : 0.0;  // inactive chains stay 0
      E[c] = act[c] ? rot_exp[tri + mm[c]] : 0;
      nxt[c] = 0.0;
      twom[c] = 2.0 * mm[c];
#pragma unroll
      for (int g = 0; g < G; ++g) accE[c][g] = accO[c][g] = 0.0;
    }
    if (any_rot) {
      // step s = l - m' contributes d_{m',m} a'_{m'}: the component of
      // parity m + s (Re for even m + s), into accE for even s, accO for odd
      // s; d_{m'-1} = d_m' (2m/B) - d_{m'+1} (A/B)
      auto step = [&](int mq, int comp, double (&acc)[C][G]) {
        double x[G];
#pragma unroll
        for (int g = 0; g < G; ++g) x[g] = xs[(2 * mq + comp) * G + g];
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
          for (int g = 0; g < G; ++g) acc[c][g] = fma(cur[c], x[g], acc[c][g]);
        const double b = ib[mq], a = ab[mq];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const double nw = fma(cur[c], twom[c] * b, -nxt[c] * a);
          nxt[c] = cur[c];
          cur[c] = nw;
        }
      };
      int mp = ell;
      while (mp >= 2) {
        const int pairs = min(4, mp >> 1);
        for (int q = 0; q < pairs; ++q) {
          step(mp, p, accE);
          step(mp - 1, p ^ 1, accO);
          mp -= 2;
        }
        // at most 8 steps since the last check: growth per step is below 3l,
        // so values stay below 2^400 * (3l)^8 < 2^1023
#pragma unroll
        for (int c = 0; c < C; ++c)
          if (fmax(fabs(cur[c]), fabs(nxt[c])) > kScaleHi) {
            cur[c] *= kScaleLo;
            nxt[c] *= kScaleLo;
            E[c] += 400;
#pragma unroll
            for (int g = 0; g < G; ++g) {
              accE[c][g] *= kScaleLo;
              accO[c][g] *= kScaleLo;
            }
          }
      }
      if (mp == 1) step(1, p, accE);
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (!act[c]) continue;
      const double d0 = ldexp(cur[c], E[c]);  // d_{0,m}
      const int64_t idx = tri + mm[c];          // l-major
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int md = mode[g];
        if (md == 0) continue;
        const int64_t o = ((int64_t)(s0 + g) * Nlm + idx) * 2;
        double re = 0.0, im = 0.0;
        if (md >= 2) {
          re = a_equ[o];
          im = a_equ[o + 1];
        }
        if (md == 3) {
          // even m: Re collects the even steps; odd m: the odd steps
          const double sr = ldexp(p ? accO[c][g] : accE[c][g], E[c]);
          const double si = ldexp(p ? accE[c][g] : accO[c][g], E[c]);
          re += fma(d0, xs[g], 2.0 * sr);
          im += fma(d0, xs[G + g], 2.0 * si);
        }
        alm[o] = re;
        alm[o + 1] = im;
      }
    }
  }
}

// a_pol -> a_tmp and a_equ -> a_pol in l-major order; the rotation and
// combine in place in a_pol; back to the healpy order into alm.
template <int G, int C>
static int launch_rotate_sets(Plan& pl, int nsets, const int* flags, int flag_stride, double* alm,
                              int* status, hipStream_t s) {
  Workspace& ws = pl.ws;
  if (!ws.a_tmp || nsets > ws.cap_sets) {
    set_error("dsht: batched rotation without its workspace");
    return DSHT_EINVAL;
  }
  const int nt = (pl.Nl + 31) / 32;
  const dim3 tg(nt, nt, nsets);
  const dim3 tgg((pl.Nl + 15) / 16, (pl.Nl + 7) / 8, (nsets + G - 1) / G);
  k_tri_group<G><<<tgg, 256, 0, s>>>((const double2*)ws.a_pol, ws.a_tmp, pl.Nl, nsets);
  DSHT_LAUNCH_CHECK();
  k_tri_transpose<true><<<tg, 256, 0, s>>>((const double2*)ws.a_equ, (double2*)ws.a_pol, pl.Nl);
  DSHT_LAUNCH_CHECK();
  dim3 grid(pl.Nl, (nsets + G - 1) / G);
  k_rotate_sets<G, C><<<grid, 256, 0, s>>>(ws.a_pol, ws.a_tmp, ws.a_pol, pl.Nl, nsets, flags,
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
  // G x C x 2 FP64 accumulators per thread: 16 sets x 2 chains, 8 x 4, 4 x 4
  if (nsets > 8) return launch_rotate_sets<16, 2>(pl, nsets, flags, flag_stride, alm, status, s);
  if (nsets > 4) return launch_rotate_sets<8, 4>(pl, nsets, flags, flag_stride, alm, status, s);
  return launch_rotate_sets<4, 4>(pl, nsets, flags, flag_stride, alm, status, s);
}

}  // namespace dsht
