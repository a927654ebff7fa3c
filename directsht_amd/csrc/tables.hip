// Stage (a): Y_lm(x_j, 0) and dY_lm/dx spline-node tables, built by recurrence
// in FP64 on the GPU.
//
// One thread per (m, node n) column runs the reference's recurrences over l:
//   P_00 = 1, P_10 = x                                 sht/legendre_py.py:19-20
//   P_mm = -sqrt(1 - 1/(2m)) sqrt(1-x^2) P_{m-1,m-1}   sht/legendre_py.py:28-30
//   P_{m+1,m} = sqrt(2m+1) x P_mm                      sht/legendre_py.py:32-34
//   P_lm = [(2l-1) x P_{l-1,m} - (l+m-1) P_{l-2,m} f2] f1/(l-m)
//                                                      sht/legendre_py.py:91-100
//   dP/dx (m = 0) = l/(1-x^2) (P_{l-1} - x P_l)        sht/legendre_py.py:54-56
//   dP/dx (m > 0) = [(l+m) f P_{l-1,m} - l x P_lm]/(1-x^2)
//                                                      sht/legendre_py.py:111-117
//   then both times sqrt((2l+1)/4pi)                   sht/legendre_py.py:77-81
// with floating-point contraction disabled and the reference's operation
// order, so with correctly rounded division and sqrt the tables equal the
// reference's numpy tables bit for bit.
//
// The per-column recurrence is inherently sequential in l; the parallelism is
// over Nl * Nx columns.  Output goes to the contraction layout (GroupDesc in
// common.h), either Y only (FP64 or FP32: {Y(x_2np), Y(x_2np+1)} per row and
// node pair, the pairs 2c and 2c+1 of a row adjacent; the contraction derives
// the Y' terms from neighbouring rows) or
// Y and Y' (FP32, {Y(2np), Y(2np+1), Y'(2np), Y'(2np+1)}: the matrix-core
// contraction), or, for DirectSHT.Yv/Yd, to the healpy row layout [Nlm][Nx].
#include "common.h"
#include "plan.h"

namespace dsht {

enum TableOut { kOutY64 = 0, kOutY32 = 1, kOutFull32 = 2, kOutHP = 3 };

template <int kOut>
__global__ __launch_bounds__(256) void k_tables(int Nl, int Nx, int Npair, double dx,
                                                const int64_t* __restrict__ mp_off,
                                                double* __restrict__ tab64,
                                                float* __restrict__ tab32,
                                                double* __restrict__ hpYv,
                                                double* __restrict__ hpYd) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (n >= 2 * Npair || m >= Nl) return;
  const bool phantom = n >= Nx;  // padding node of an odd Nx: zeros
  const double x = phantom ? 0.0 : (double)n * dx;
  const double sx = sqrt(1.0 - x * x);
  const double omx2 = 1.0 - x * x;
  const int np = n >> 1, odd = n & 1;

  auto emit = [&](int ell, double P, double D) {
    const double norm = sqrt((double)(2 * ell + 1) / 4.0 / M_PI);
    double y = phantom ? 0.0 : P * norm;
    double yd = phantom ? 0.0 : D * norm;
    if constexpr (kOut == kOutHP) {
      if (!phantom) {
        int64_t row = hp_index(ell, m, Nl);
        hpYv[row * Nx + n] = y;
        hpYd[row * Nx + n] = yd;
      }
    } else {
      const int k = ell - m, P2 = k & 1, i = k >> 1;
      const int g = i / kGroupRows, r = i % kGroupRows;
      const int R = rows_mp(Nl, m, P2);
      const int nrows = min(kGroupRows, R - g * kGroupRows);
      const int64_t gbase = mp_off[2 * m + P2] + (int64_t)g * kGroupRows * Npair;
      const int64_t e = gbase + (int64_t)np * nrows + r;
      // Y-only: the two pairs 2c, 2c+1 of a row side by side,
      // [c][row][pair parity][node parity] (Npair is even)
      const int64_t ey = gbase + (int64_t)(np >> 1) * 2 * nrows + 2 * r + (np & 1);
      if constexpr (kOut == kOutY64) tab64[ey * 2 + odd] = y;
      if constexpr (kOut == kOutY32) tab32[ey * 2 + odd] = (float)y;
      if constexpr (kOut == kOutFull32) {
        tab32[e * 4 + odd] = (float)y;
        tab32[e * 4 + 2 + odd] = (float)yd;
      }
    }
  };

  // P_mm by the chain over m' = 1..m (each column recomputes its own diagonal)
  double pmm = 1.0;
  for (int k = 1; k <= m; ++k) pmm = (-sqrt(1.0 - 1.0 / (double)(2 * k)) * sx) * pmm;

  double p2 = 0.0, p1 = 0.0;  // P_{l-2,m}, P_{l-1,m}
  for (int ell = m; ell < Nl; ++ell) {
    double p;
    if (ell == m) {
      p = pmm;
    } else if (ell == m + 1) {
      p = (m == 0) ? x : (sqrt((double)(2 * m) + 1.0) * x) * pmm;
    } else {
      const double f1 = sqrt((double)(ell - m) * 1.0 / (double)(ell + m));
      const double f2 = sqrt(((double)(ell - m) - 1.0) / ((double)(ell + m) - 1.0));
      const double a = ((double)(2 * ell - 1) * x) * p1;
      const double b = ((double)(ell + m - 1) * p2) * f2;
      p = (a - b) * (f1 / (double)(ell - m));
    }
    double d;
    if (m == 0) {
      if (ell == 0) d = 0.0;
      else if (ell == 1) d = 1.0;
      else d = ((double)ell / (1.0 - x * x)) * (p1 - x * p);
    } else {
      const double f = sqrt((double)(ell - m) / (double)(ell + m));
      const double lhs = ((double)(ell + m) * f) * (ell == m ? 0.0 : p1);
      d = (lhs - ((double)ell * x) * p) / omx2;
    }
    emit(ell, p, d);
    p2 = p1;
    p1 = p;
  }
}

// (l, m) of healpy row `row`: inverts row = m(2Nl-1-m)/2 + l, as
// sht/shared_utils.py:93-98 (the estimate corrected to the exact m)
__device__ inline void hp_row_lm(int64_t row, int Nl, int& ell, int& m) {
  const double b = 2.0 * Nl - 1.0;
  m = (int)ceil((b - sqrt(b * b - 8.0 * ((double)row - (Nl - 1)))) / 2.0);
  if (m < 0) m = 0;
  while (m > 0 && hp_index(m, m, Nl) > row) --m;
  while (m + 1 < Nl && hp_index(m + 1, m + 1, Nl) <= row) ++m;
  ell = (int)(row - hp_index(0, m, Nl));
}

// Copy caller-provided healpy-row tables (csht's make_table layout) into the
// contraction layout (FP64).  One block per healpy row.
__global__ __launch_bounds__(256) void k_relayout(int Nl, int Nx, int Npair,
                                                  const int64_t* __restrict__ mp_off,
                                                  const double* __restrict__ hpYv,
                                                  const double* __restrict__ hpYd,
                                                  double* __restrict__ tab64) {
  const int64_t row = blockIdx.x;
  int ell, m;
  hp_row_lm(row, Nl, ell, m);
  const int k = ell - m, P = k & 1, i = k >> 1;
  const int g = i / kGroupRows, r = i % kGroupRows;
  const int nrows = min(kGroupRows, rows_mp(Nl, m, P) - g * kGroupRows);
  const int64_t base = mp_off[2 * m + P] + (int64_t)g * kGroupRows * Npair + r;
  for (int n = threadIdx.x; n < 2 * Npair; n += blockDim.x) {
    const int64_t idx4 = base + (int64_t)(n >> 1) * nrows;
    const double y = n < Nx ? hpYv[row * Nx + n] : 0.0;
    const double yd = n < Nx ? hpYd[row * Nx + n] : 0.0;
    tab64[idx4 * 4 + (n & 1)] = y;
    tab64[idx4 * 4 + 2 + (n & 1)] = yd;
  }
}

// Healpy rows row0 .. row0 + gridDim.x - 1 of a Y table (`slab`: those rows,
// [row][Nx], FP64) into the plan's Y-only table (FP64 or FP32), at the
// positions k_tables writes: a plan built from a table saved earlier.  One
// block per row.
template <typename T>
__global__ __launch_bounds__(256) void k_relayout_y(int Nl, int Nx, int Npair,
                                                    const int64_t* __restrict__ mp_off,
                                                    int64_t row0, const double* __restrict__ slab,
                                                    T* __restrict__ tab) {
  const int64_t row = row0 + blockIdx.x;
  int ell, m;
  hp_row_lm(row, Nl, ell, m);
  const int k = ell - m, P = k & 1, i = k >> 1;
  const int g = i / kGroupRows, r = i % kGroupRows;
  const int nrows = min(kGroupRows, rows_mp(Nl, m, P) - g * kGroupRows);
  const int64_t gbase = mp_off[2 * m + P] + (int64_t)g * kGroupRows * Npair;
  const double* src = slab + (int64_t)blockIdx.x * Nx;
  for (int n = threadIdx.x; n < 2 * Npair; n += blockDim.x) {
    const int np = n >> 1;
    const int64_t ey = gbase + (int64_t)(np >> 1) * 2 * nrows + 2 * r + (np & 1);
    tab[ey * 2 + (n & 1)] = (T)(n < Nx ? src[n] : 0.0);  // the padding node of an odd Nx: 0
  }
}

int launch_relayout_y(const Plan& pl, int64_t row0, int64_t nrows, const double* slab,
                      hipStream_t s) {
  if (nrows <= 0 || pl.otf) return DSHT_OK;  // FP32 plans hold no Y table
  if (pl.precision == DSHT_FP64)
    k_relayout_y<double><<<(unsigned)nrows, 256, 0, s>>>(pl.Nl, pl.Nx, pl.Npair, pl.mp_off, row0,
                                                         slab, pl.tab64y);
  else
    k_relayout_y<float><<<(unsigned)nrows, 256, 0, s>>>(pl.Nl, pl.Nx, pl.Npair, pl.mp_off, row0,
                                                        slab, pl.tab32y);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_relayout(const Plan& pl, const double* hpYv, const double* hpYd, hipStream_t s) {
  k_relayout<<<(unsigned)pl.Nlm, 256, 0, s>>>(pl.Nl, pl.Nx, pl.Npair, pl.mp_off, hpYv, hpYd, pl.tab64);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_tables(const Plan& pl, int out, void* dst, double* hpYd, hipStream_t s) {
  dim3 block(256);
  dim3 grid((2 * pl.Npair + 255) / 256, pl.Nl);
  double* d64 = static_cast<double*>(dst);
  float* d32 = static_cast<float*>(dst);
  switch (out) {
    case kOutY64:
      k_tables<kOutY64><<<grid, block, 0, s>>>(pl.Nl, pl.Nx, pl.Npair, pl.dx, pl.mp_off, d64, nullptr,
                                               nullptr, nullptr);
      break;
    case kOutY32:
      k_tables<kOutY32><<<grid, block, 0, s>>>(pl.Nl, pl.Nx, pl.Npair, pl.dx, pl.mp_off, nullptr, d32,
                                               nullptr, nullptr);
      break;
    case kOutFull32:
      k_tables<kOutFull32><<<grid, block, 0, s>>>(pl.Nl, pl.Nx, pl.Npair, pl.dx, pl.mp_off, nullptr,
                                                  d32, nullptr, nullptr);
      break;
    default:
      k_tables<kOutHP><<<grid, block, 0, s>>>(pl.Nl, pl.Nx, pl.Npair, pl.dx, pl.mp_off, nullptr,
                                              nullptr, d64, hpYd);
  }
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
