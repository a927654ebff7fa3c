// Harmonic synthesis on iso-latitude rings (the inverse of the transform's
// sum over points): for alm in healpy order and ring cosines x_r,
//     F[r][m] = sum_{l=m}^{Nl-1} a_lm Ybar_lm(x_r),
// Ybar the fully normalized associated Legendre function with the
// Condon-Shortley phase, so that a real field on the ring is
//     f(x_r, phi) = Re F[r][0] + 2 Re sum_{m>0} F[r][m] e^{i m phi}
// (one inverse real FFT per ring).  This is what healpy's synfast does for
// the reference's lognormal mocks (sht/mocks.py:60); here it feeds the
// mock generator on the device.
//
// Block (m, 1024 rings), four rings per thread.  The l-recurrence
// coefficients of order m are built once per block in LDS:
//     Ybar_mm = (-1)^m c_m sin^m,  c_m = prod_{k<=m} sqrt((2k+1)/(2k)) / sqrt(4 pi),
//     Ybar_{m+1,m} = sqrt(2m+3) x Ybar_mm,
//     Ybar_lm = a_l (x Ybar_{l-1,m} - b_l Ybar_{l-2,m}),
//     a_l = sqrt((4l^2-1)/(l^2-m^2)),  b_l = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1));
// each thread walks l for its four rings (independent recurrences), the
// a_lm being uniform (scalar) loads.
// sin^m underflows to zero only where Ybar_lm is below ~1e-300 for every
// l < Nl <= 2048.
#include "common.h"
#include "plan.h"

namespace dsht {

constexpr int kSynthMaxNl = 2048;
constexpr int kSynthR = 4;  // rings per thread

__global__ __launch_bounds__(256) void k_synth_rings(const double2* __restrict__ alm, int Nl,
                                                     int nrings, const double* __restrict__ x,
                                                     double2* __restrict__ F) {
  __shared__ double sa[kSynthMaxNl], sb[kSynthMaxNl];
  __shared__ double s_cm;
  const int m = Nl - 1 - (int)blockIdx.x;  // longest sums first
  for (int l = m + 2 + (int)threadIdx.x; l < Nl; l += blockDim.x) {
    const double L = l, M = m;
    sa[l] = sqrt((4.0 * L * L - 1.0) / (L * L - M * M));
    sb[l] = sqrt(((L - 1.0) * (L - 1.0) - M * M) / (4.0 * (L - 1.0) * (L - 1.0) - 1.0));
  }
  if (threadIdx.x == 0) {
    double c = 0.28209479177387814;  // 1/sqrt(4 pi)
    for (int k = 1; k <= m; ++k) c *= sqrt((2.0 * k + 1.0) / (2.0 * k));
    s_cm = c;
  }
  __syncthreads();
  // kSynthR rings per thread (r, r + blockDim, ...), independent recurrences
  const int rb = blockIdx.y * (kSynthR * blockDim.x) + threadIdx.x;
  if (rb >= nrings) return;
  const double2* a = alm + hp_index(m, m, Nl);  // a_lm, l = m .. Nl-1, contiguous
  const double cm = m & 1 ? -s_cm : s_cm;
  double xr[kSynthR], p1[kSynthR], p2[kSynthR], fr[kSynthR], fi[kSynthR];
#pragma unroll
  for (int k = 0; k < kSynthR; ++k) {
    const int r = rb + k * (int)blockDim.x;
    xr[k] = r < nrings ? x[r] : 0.0;
    const double sr = sqrt(fmax(0.0, (1.0 - xr[k]) * (1.0 + xr[k])));
    p2[k] = cm * (m > 0 ? pow(sr, (double)m) : 1.0);  // Ybar_mm
    fr[k] = a[0].x * p2[k];
    fi[k] = a[0].y * p2[k];
  }
  if (m + 1 < Nl) {
    const double c1 = sqrt(2.0 * m + 3.0);
#pragma unroll
    for (int k = 0; k < kSynthR; ++k) {
      p1[k] = c1 * xr[k] * p2[k];  // Ybar_{m+1,m}
      fr[k] = fma(a[1].x, p1[k], fr[k]);
      fi[k] = fma(a[1].y, p1[k], fi[k]);
    }
    for (int l = m + 2; l < Nl; ++l) {
      const double al = sa[l], bl = sb[l];
      const double2 c = a[l - m];
#pragma unroll
      for (int k = 0; k < kSynthR; ++k) {
        const double p = al * fma(xr[k], p1[k], -bl * p2[k]);
        fr[k] = fma(c.x, p, fr[k]);
        fi[k] = fma(c.y, p, fi[k]);
        p2[k] = p1[k];
        p1[k] = p;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kSynthR; ++k) {
    const int r = rb + k * (int)blockDim.x;
    if (r < nrings) F[(int64_t)r * Nl + m] = make_double2(fr[k], fi[k]);
  }
}

int launch_synth_rings(int Nl, const double* alm, int nrings, const double* x, double* F,
                       hipStream_t s) {
  dim3 grid(Nl, (nrings + 256 * kSynthR - 1) / (256 * kSynthR));
  k_synth_rings<<<grid, 256, 0, s>>>(reinterpret_cast<const double2*>(alm), Nl, nrings, x,
                                     reinterpret_cast<double2*>(F));
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int synth_max_nl() { return kSynthMaxNl; }

// grid.y = ceil(nrings / (256 kSynthR)) must stay within 65535
int synth_max_rings() { return 65535 * 256 * kSynthR; }

}  // namespace dsht
