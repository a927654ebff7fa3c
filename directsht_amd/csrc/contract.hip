// Stage (c): contraction of the resident tables with the node coefficients,
//   a_lm = sum_n Y_lm(x_n) c_P(m, n) + Y'_lm(x_n) d_P(m, n),  P = (l-m) & 1,
// for both branches (equatorial, polar) at once -- the per-(l,m) dot products
// of sht/interp_funcs_py.py:46-60 / sht/interp_funcs_jax.py:31-48 and
// csht/sht_helper.c:255-274, as one HBM-streaming GEMV.
//
// One wavefront per row group (up to 64 rows l of one (m, P)), launched in an
// XCD-aware order (xcd_schedule in api.hip); lane r owns row
// l = m + P + 2(i0 + r) and walks the node pairs.  Each step loads one
// 4-vector {Y(x_2np), Y(x_2np+1), Y'(x_2np), Y'(x_2np+1)} per lane (a
// contiguous 1 KB (FP32) / 2 KB (FP64) block per wavefront) and multiplies it
// with the wave-uniform node coefficients of that pair, which the compiler
// keeps in scalar registers.  FP32 tables accumulate 16 node pairs in FP32 and
// then in FP64; FP64 tables accumulate in FP64.  The result goes straight to
// the healpy index of (l, m) (stage (d) pack).
#include "common.h"
#include "plan.h"

namespace dsht {

constexpr int kInner = 16;  // node pairs per FP32 partial sum

// Table entries are read exactly once per transform: non-temporal loads keep
// the stream from displacing the node coefficients (reused by every row
// group of an (m, P)) in the caches.
__device__ inline float4 stream_load(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline double4 stream_load(const double4* p) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_double4(v.x, v.y, v.z, v.w);
}

__global__ __launch_bounds__(64) void k_contract32(const float4* __restrict__ tab,
                                                   const float4* __restrict__ rhs,
                                                   const GroupDesc* __restrict__ groups, int Npair,
                                                   int Nl, int nsets, int set,
                                                   double* __restrict__ a_equ,
                                                   double* __restrict__ a_pol) {
  const GroupDesc gd = groups[blockIdx.x];
  if (gd.nrows == 0) return;  // padding of the XCD schedule
  const int r = threadIdx.x;
  const bool valid = r < gd.nrows;
  const float4* t = tab + gd.off + r;
  // the set's four coefficient vectors of node pair np start at q[np * qs]
  const float4* q = rhs + ((int64_t)(gd.m * 2 + gd.P) * Npair * nsets + set) * 4;
  const int64_t qs = 4 * (int64_t)nsets;
  const int64_t stride = gd.nrows;
  double aer = 0, aei = 0, apr = 0, api = 0;
  for (int np0 = 0; np0 < Npair; np0 += kInner) {
    float fer = 0.f, fei = 0.f, fpr = 0.f, fpi = 0.f;
    const int np1 = min(Npair, np0 + kInner);
#pragma unroll 8
    for (int np = np0; np < np1; ++np) {
      const float4 y = valid ? stream_load(t + np * stride) : make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 qer = q[np * qs + 0], qei = q[np * qs + 1];
      const float4 qpr = q[np * qs + 2], qpi = q[np * qs + 3];
      fer = fmaf(y.x, qer.x, fmaf(y.y, qer.y, fmaf(y.z, qer.z, fmaf(y.w, qer.w, fer))));
      fei = fmaf(y.x, qei.x, fmaf(y.y, qei.y, fmaf(y.z, qei.z, fmaf(y.w, qei.w, fei))));
      fpr = fmaf(y.x, qpr.x, fmaf(y.y, qpr.y, fmaf(y.z, qpr.z, fmaf(y.w, qpr.w, fpr))));
      fpi = fmaf(y.x, qpi.x, fmaf(y.y, qpi.y, fmaf(y.z, qpi.z, fmaf(y.w, qpi.w, fpi))));
    }
    aer += fer;
    aei += fei;
    apr += fpr;
    api += fpi;
  }
  if (valid) {
    const int ell = gd.m + gd.P + 2 * (gd.i0 + r);
    const int64_t idx = hp_index(ell, gd.m, Nl) + (int64_t)set * Nl * (Nl + 1) / 2;
    a_equ[2 * idx] = aer;
    a_equ[2 * idx + 1] = aei;
    a_pol[2 * idx] = apr;
    a_pol[2 * idx + 1] = api;
  }
}

__global__ __launch_bounds__(64) void k_contract64(const double4* __restrict__ tab,
                                                   const double4* __restrict__ rhs,
                                                   const GroupDesc* __restrict__ groups, int Npair,
                                                   int Nl, int nsets, int set,
                                                   double* __restrict__ a_equ,
                                                   double* __restrict__ a_pol) {
  const GroupDesc gd = groups[blockIdx.x];
  if (gd.nrows == 0) return;  // padding of the XCD schedule
  const int r = threadIdx.x;
  const bool valid = r < gd.nrows;
  const double4* t = tab + gd.off + r;
  // the set's four coefficient vectors of node pair np start at q[np * qs]
  const double4* q = rhs + ((int64_t)(gd.m * 2 + gd.P) * Npair * nsets + set) * 4;
  const int64_t qs = 4 * (int64_t)nsets;
  const int64_t stride = gd.nrows;
  double aer = 0, aei = 0, apr = 0, api = 0;
#pragma unroll 4
  for (int np = 0; np < Npair; ++np) {
    const double4 y = valid ? stream_load(t + np * stride) : make_double4(0, 0, 0, 0);
    const double4 qer = q[np * qs + 0], qei = q[np * qs + 1];
    const double4 qpr = q[np * qs + 2], qpi = q[np * qs + 3];
    aer = fma(y.x, qer.x, fma(y.y, qer.y, fma(y.z, qer.z, fma(y.w, qer.w, aer))));
    aei = fma(y.x, qei.x, fma(y.y, qei.y, fma(y.z, qei.z, fma(y.w, qei.w, aei))));
    apr = fma(y.x, qpr.x, fma(y.y, qpr.y, fma(y.z, qpr.z, fma(y.w, qpr.w, apr))));
    api = fma(y.x, qpi.x, fma(y.y, qpi.y, fma(y.z, qpi.z, fma(y.w, qpi.w, api))));
  }
  if (valid) {
    const int ell = gd.m + gd.P + 2 * (gd.i0 + r);
    const int64_t idx = hp_index(ell, gd.m, Nl) + (int64_t)set * Nl * (Nl + 1) / 2;
    a_equ[2 * idx] = aer;
    a_equ[2 * idx + 1] = aei;
    a_pol[2 * idx] = apr;
    a_pol[2 * idx + 1] = api;
  }
}

int launch_contract(Plan& pl, int nsets, hipStream_t s) {
  if (pl.precision == DSHT_FP32 && nsets >= kMfmaMinSets && pl.contract == DSHT_CONTRACT_MATRIX)
    return launch_contract_mfma(pl, nsets, s);
  for (int set = 0; set < nsets; ++set) {
    if (pl.precision == DSHT_FP32) {
      k_contract32<<<pl.ngroups, 64, 0, s>>>(reinterpret_cast<const float4*>(pl.tab32),
                                             reinterpret_cast<const float4*>(pl.ws.rhs), pl.groups,
                                             pl.Npair, pl.Nl, nsets, set, pl.ws.a_equ, pl.ws.a_pol);
    } else {
      k_contract64<<<pl.ngroups, 64, 0, s>>>(reinterpret_cast<const double4*>(pl.tab64),
                                             reinterpret_cast<const double4*>(pl.ws.rhs), pl.groups,
                                             pl.Npair, pl.Nl, nsets, set, pl.ws.a_equ, pl.ws.a_pol);
    }
    DSHT_LAUNCH_CHECK();
  }
  return DSHT_OK;
}

}  // namespace dsht
