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
// Block (m, 256 rings).  The l-recurrence coefficients of order m are built
// once per block in LDS:
//     Ybar_mm = (-1)^m c_m sin^m,  c_m = prod_{k<=m} sqrt((2k+1)/(2k)) / sqrt(4 pi),
//     Ybar_{m+1,m} = sqrt(2m+3) x Ybar_mm,
//     Ybar_lm = a_l (x Ybar_{l-1,m} - b_l Ybar_{l-2,m}),
//     a_l = sqrt((4l^2-1)/(l^2-m^2)),  b_l = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1));
// each thread walks l for its ring, the a_lm being uniform (scalar) loads.
// sin^m underflows to zero only where Ybar_lm is below ~1e-300 for every
// l < Nl <= 2048.
#include "common.h"
#include "plan.h"

namespace dsht {

constexpr int kSynthMaxNl = 2048;

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
  const int r = blockIdx.y * blockDim.x + threadIdx.x;
  if (r >= nrings) return;
  const double xr = x[r];
  const double sr = sqrt(fmax(0.0, (1.0 - xr) * (1.0 + xr)));
  const double2* a = alm + hp_index(m, m, Nl);  // a_lm, l = m .. Nl-1, contiguous
  double p2 = (m & 1 ? -s_cm : s_cm) * (m > 0 ? pow(sr, (double)m) : 1.0);  // Ybar_mm
  double fre = a[0].x * p2, fim = a[0].y * p2;
  if (m + 1 < Nl) {
    double p1 = sqrt(2.0 * m + 3.0) * xr * p2;  // Ybar_{m+1,m}
    fre = fma(a[1].x, p1, fre);
    fim = fma(a[1].y, p1, fim);
    for (int l = m + 2; l < Nl; ++l) {
      const double p = sa[l] * fma(xr, p1, -sb[l] * p2);
      const double2 c = a[l - m];
      fre = fma(c.x, p, fre);
      fim = fma(c.y, p, fim);
      p2 = p1;
      p1 = p;
    }
  }
  F[(int64_t)r * Nl + m] = make_double2(fre, fim);
}

int launch_synth_rings(int Nl, const double* alm, int nrings, const double* x, double* F,
                       hipStream_t s) {
  dim3 grid(Nl, (nrings + 255) / 256);
  k_synth_rings<<<grid, 256, 0, s>>>(reinterpret_cast<const double2*>(alm), Nl, nrings, x,
                                     reinterpret_cast<double2*>(F));
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int synth_max_nl() { return kSynthMaxNl; }

}  // namespace dsht
