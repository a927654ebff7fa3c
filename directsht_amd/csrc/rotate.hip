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
#include "common.h"
#include "plan.h"

namespace dsht {

constexpr double kScaleHi = 0x1p400;
constexpr double kScaleLo = 0x1p-400;

__global__ __launch_bounds__(256) void k_rotate_combine(const double* __restrict__ a_equ,
                                                        const double* __restrict__ a_pol,
                                                        double* __restrict__ alm, int Nl,
                                                        const int* __restrict__ flags,
                                                        int* __restrict__ status,
                                                        const double* __restrict__ rot_top,
                                                        const int* __restrict__ rot_exp,
                                                        const double* __restrict__ rot_AB,
                                                        const double* __restrict__ rot_iB) {
  const int ell = blockIdx.x;
  const int fl = flags[0];
  const int npol = flags[1];
  if (status && ell == 0 && threadIdx.x == 0) *status = fl;
  if (fl & DSHT_STATUS_POLAR_RANGE) {
    for (int m = threadIdx.x; m <= ell; m += blockDim.x) {
      const int64_t idx = hp_index(ell, m, Nl);
      alm[2 * idx] = 0.0;
      alm[2 * idx + 1] = 0.0;
    }
    return;
  }
  if (npol == 0) {
    for (int m = threadIdx.x; m <= ell; m += blockDim.x) {
      const int64_t idx = hp_index(ell, m, Nl);
      alm[2 * idx] = a_equ[2 * idx];
      alm[2 * idx + 1] = a_equ[2 * idx + 1];
    }
    return;
  }
  extern __shared__ double lds[];
  double* sre = lds;                  // Re a'_{l m'}
  double* sim = lds + (ell + 1);      // Im a'_{l m'}
  double* sAB = lds + 2 * (ell + 1);  // A(m')/B(m')
  double* sB = lds + 3 * (ell + 1);   // 1/B(m')
  const int64_t tri = (int64_t)ell * (ell + 1) / 2;
  for (int mp = threadIdx.x; mp <= ell; mp += blockDim.x) {
    const int64_t idx = hp_index(ell, mp, Nl);
    sre[mp] = a_pol[2 * idx];
    sim[mp] = a_pol[2 * idx + 1];
    sAB[mp] = rot_AB[tri + mp];
    sB[mp] = rot_iB[tri + mp];
  }
  __syncthreads();
  for (int m = threadIdx.x; m <= ell; m += blockDim.x) {
    // Step s = l - m' contributes d_{m',m} * Re a'_{m'} when l + m + m' is
    // even and d_{m',m} * Im a'_{m'} when it is odd, i.e. by the parity of
    // m + s: even steps use xe, odd steps xo (no divergent branch).
    const double* xe = (m & 1) ? sim : sre;
    const double* xo = (m & 1) ? sre : sim;
    const double twom = 2.0 * m;
    double cur = rot_top[tri + m], nxt = 0.0;
    int E = rot_exp[tri + m];
    double accE = 0.0, accO = 0.0;
    int mp = ell;
    // d_{mp-1} = d_mp * (2m/B) - d_{mp+1} * (A/B): one dependent FMA per step
    while (mp >= 2) {
      const int pairs = min(4, mp >> 1);
      for (int q = 0; q < pairs; ++q) {
        accE = fma(ldexp(cur, E), xe[mp], accE);
        double nw = fma(cur, twom * sB[mp], -nxt * sAB[mp]);
        nxt = cur;
        cur = nw;
        --mp;
        accO = fma(ldexp(cur, E), xo[mp], accO);
        nw = fma(cur, twom * sB[mp], -nxt * sAB[mp]);
        nxt = cur;
        cur = nw;
        --mp;
      }
      // at most 8 steps since the last check: growth per step is below 3l,
      // so values stay below 2^400 * (3l)^8 < 2^1023
      if (fmax(fabs(cur), fabs(nxt)) > kScaleHi) {
        cur *= kScaleLo;
        nxt *= kScaleLo;
        E += 400;
      }
    }
    if (mp == 1) {
      accE = fma(ldexp(cur, E), xe[1], accE);
      cur = fma(cur, twom * sB[1], -nxt * sAB[1]);
    }
    const double d0 = ldexp(cur, E);  // d_{0,m}
    const double sr = (m & 1) ? accO : accE;
    const double si = (m & 1) ? accE : accO;
    const int64_t idx = hp_index(ell, m, Nl);
    alm[2 * idx] = a_equ[2 * idx] + fma(d0, sre[0], 2.0 * sr);
    alm[2 * idx + 1] = a_equ[2 * idx + 1] + fma(d0, sim[0], 2.0 * si);
  }
}

int launch_rotate_combine(Plan& pl, double* alm, int* status, int mode, hipStream_t s) {
  (void)mode;
  const size_t lds = sizeof(double) * 4 * (size_t)pl.Nl;
  k_rotate_combine<<<pl.Nl, 256, lds, s>>>(pl.ws.a_equ, pl.ws.a_pol, alm, pl.Nl, pl.ws.flags, status,
                                           pl.rot_top, pl.rot_exp, pl.rot_AB, pl.rot_iB);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
