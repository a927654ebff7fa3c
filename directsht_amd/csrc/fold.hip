// Stage (b), part 3: reduce the per-run partials of each segment and fold the
// four Hermite components onto the spline nodes and the two hemispheres.
//
// The reference contracts per bin j (sht/interp_funcs_py.py:46-60):
//   a = sum_j Y(x_j) v0(j) + dx Y'(x_j) v1(j) + Y(x_{j+1}) v2(j) + dx Y'(x_{j+1}) v3(j)
// and adds the x < 0 hemisphere with (-1)^(l-m) (sht/sht.py:185,279).  The
// same sum, regrouped by node n, is
//   a = sum_n Y(x_n) c(n) + Y'(x_n) d(n),
//   c(n) = v0(n) + v2(n-1),   d(n) = dx (v1(n) + v3(n-1)),
// and with s = c+ + c-, delta = c+ - c- the rows with even l-m contract s and
// the rows with odd l-m contract delta -- one table read for both hemispheres.
//
// One block per (node n, branch); threads over m.  The runs of a segment are
// contiguous in run order (runs are numbered by position in the sorted
// array); they are found by binary search in run_seg and summed in order.
// Output rhs[m][P][np][16] (P = parity class, np = node pair), as the
// contraction reads it: for branch b, 8 values at offset 8b:
//   re: c(2np), c(2np+1), d(2np), d(2np+1);  im: the same four.
#include "common.h"
#include "plan.h"

namespace dsht {

__device__ inline int lower_bound_int(const int* a, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename TP, typename T>
__global__ __launch_bounds__(256) void k_fold(const TP* __restrict__ part,
                                              const int* __restrict__ run_seg,
                                              const int* __restrict__ run_total, int Nl, int Nx,
                                              int Npair, double dx, T* __restrict__ rhs) {
  const int n = blockIdx.x, br = blockIdx.y;
  const int np = n >> 1, odd = n & 1;
  __shared__ int rr[4][2];
  if (threadIdx.x < 4) {
    const int h = threadIdx.x >> 1;       // hemisphere
    const int j = n - (threadIdx.x & 1);  // bin n (v0,v1) or n-1 (v2,v3)
    int lo = 0, hi = 0;
    if (j >= 0 && j <= Nx - 2) {
      const int s = (br * 2 + h) * (Nx - 1) + j;
      const int nr = *run_total;
      lo = lower_bound_int(run_seg, nr, s);
      hi = lower_bound_int(run_seg, nr, s + 1);
    }
    rr[threadIdx.x][0] = lo;
    rr[threadIdx.x][1] = hi;
  }
  __syncthreads();
  for (int m = threadIdx.x; m < Nl; m += blockDim.x) {
    double c[2][2] = {{0, 0}, {0, 0}}, d[2][2] = {{0, 0}, {0, 0}};  // [hemisphere][re/im]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int h = q >> 1, half = q & 1;  // half 1: v2,v3 of bin n-1
      for (int r = rr[q][0]; r < rr[q][1]; ++r) {
        const TP* v = part + (((int64_t)r * 2 + half) * Nl + m) * 4;
        c[h][0] += (double)v[0];
        c[h][1] += (double)v[1];
        d[h][0] += (double)v[2];
        d[h][1] += (double)v[3];
      }
    }
#pragma unroll
    for (int P = 0; P < 2; ++P) {
      const double sg = P ? -1.0 : 1.0;
      T* o = rhs + (((int64_t)m * 2 + P) * Npair + np) * 16 + 8 * br;
      o[odd] = (T)(c[0][0] + sg * c[1][0]);
      o[2 + odd] = (T)(dx * (d[0][0] + sg * d[1][0]));
      o[4 + odd] = (T)(c[0][1] + sg * c[1][1]);
      o[6 + odd] = (T)(dx * (d[0][1] + sg * d[1][1]));
    }
  }
}

int launch_fold(Plan& pl, hipStream_t s) {
  dim3 grid(2 * pl.Npair, 2);
  const int* run_total = pl.ws.run_base + pl.ws.cur_chunks;  // scan total = number of runs
  if (pl.precision == DSHT_FP32) {
    k_fold<float, float><<<grid, 256, 0, s>>>((const float*)pl.ws.part, pl.ws.run_seg, run_total,
                                              pl.Nl, pl.Nx, pl.Npair, pl.dx, (float*)pl.ws.rhs);
  } else {
    k_fold<double, double><<<grid, 256, 0, s>>>((const double*)pl.ws.part, pl.ws.run_seg, run_total,
                                                pl.Nl, pl.Nx, pl.Npair, pl.dx, (double*)pl.ws.rhs);
  }
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
