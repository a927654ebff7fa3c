// Internal definitions shared by the HIP sources of libdsht.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/dsht.h"

namespace dsht {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

#define DSHT_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::dsht::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return DSHT_EHIP;                                                             \
    }                                                                               \
  } while (0)

#define DSHT_LAUNCH_CHECK()                                                         \
  do {                                                                              \
    hipError_t e_ = hipGetLastError();                                              \
    if (e_ != hipSuccess) {                                                         \
      ::dsht::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));    \
      return DSHT_EHIP;                                                             \
    }                                                                               \
  } while (0)

// Internal transform mode of the csht-compatible do_transform: the "theta"
// input holds cos(theta) of points of one sign, only |cos(theta)| is used and
// no parity or polar split is applied (csht/sht_helper.c:176-286).
constexpr int DSHT_MODE_CSHT = 2;

// ---------------------------------------------------------------------------
// Layout of the contraction tables
// ---------------------------------------------------------------------------
// For every m and parity P = (l - m) & 1 the rows l = m + P + 2i,
// i = 0 .. R(m,P)-1, are cut into groups of up to 64 rows (one wavefront
// each).  A group with nrows rows stores, for every node pair np
// (nodes 2np, 2np+1), nrows consecutive 4-vectors
//     { Y(x_2np), Y(x_2np+1), Y'(x_2np), Y'(x_2np+1) }
// so one wavefront load at fixed np is one contiguous nrows*16-byte (FP32) or
// nrows*32-byte (FP64) block, and a group is one contiguous stream.
constexpr int kGroupRows = 64;

struct GroupDesc {
  int m;        // order
  int P;        // parity of l - m
  int i0;       // first row index i within (m, P)
  int nrows;    // rows in this group (<= 64)
  int64_t off;  // offset of the group, in 4-vectors
};

__host__ __device__ inline int rows_mp(int Nl, int m, int P) {
  int r = Nl - m - P;
  return r > 0 ? (r + 1) / 2 : 0;
}

__host__ __device__ inline int64_t hp_index(int ell, int m, int Nl) {
  return ((int64_t)m * (2 * Nl - 1 - m)) / 2 + ell;
}

// The plain node coefficients (one set at a time): for node pair np, branch
// br (0 equatorial, 1 polar), order m and parity P, 12 values
// [c | f | g][re | im][node 2np, 2np + 1] at plain_rec(np, br, m, P, Nl).  A
// branch's records of one pair (all orders, both parities) are one block, so
// the kernels that produce them write whole lines.
__host__ __device__ inline int64_t plain_rec(int np, int br, int m, int P, int Nl) {
  return ((((int64_t)np * 2 + br) * Nl + m) * 2 + P) * 12;
}

// ---------------------------------------------------------------------------
// Segments of the sorted point array
// ---------------------------------------------------------------------------
// seg = (branch*2 + hemisphere) * (Nx-1) + bin, branch 0 = equatorial,
// 1 = polar (rotated); hemisphere 0: x >= 0, 1: x < 0; bin j in [0, Nx-2].
// Points that the reference drops (NaN theta) get seg = nseg (sorted last,
// never scattered).
// Input points per counting-sort tile: one wavefront ranks a tile's points
// in input order, so about kSortTiles tiles (a multiple of 256 points each,
// at least 256) fill the device in one round.  The ranks count in 16 bits,
// so a tile holds at most kSortTileMax points: transforms of more than
// kSortTiles * kSortTileMax (about 8.4e7) points take more tiles.
constexpr int kSortTiles = 1280;
constexpr int kSortRound = 256;  // points a wavefront takes per step (4 per lane)
constexpr int kSortTileMax = 65280;  // largest multiple of kSortRound below 2^16
__host__ __device__ inline int sort_tile(int64_t Np) {
  int64_t t = (Np + kSortTiles - 1) / kSortTiles;
  t = (t + kSortRound - 1) / kSortRound * kSortRound;
  t = t < kSortRound ? kSortRound : t;
  return (int)(t > kSortTileMax ? kSortTileMax : t);
}
// tile rows of the histogram matrix for any transform of up to Np points
inline int64_t sort_tiles_cap(int64_t Np) {
  const int64_t t = (Np + kSortRound - 1) / kSortRound;
  const int64_t big = (Np + kSortTileMax - 1) / kSortTileMax;
  const int64_t c = t < kSortTiles ? t : kSortTiles;
  return c < 1 ? 1 : (c > big ? c : big);
}

inline int num_segments(int Nx) { return 4 * (Nx - 1); }

// *p = max(*p, v) for non-negative floats kept as their bits (which order
// like the floats).  The value is read first and the atomic skipped when it
// cannot raise it: atomics on one address serialize, and once the maximum is
// in place most callers would only repeat it.  A stale read can only be
// smaller than the current value, so no update is lost.
__device__ inline void atomic_max_pos(unsigned* p, float v) {
  if (!(v > 0.f)) return;
  const unsigned cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  if (__float_as_uint(v) > cur) atomicMax(p, __float_as_uint(v));
}

// Per-point record after sorting: one 64-byte line per point, written once by
// the placement pass and read by the scatter.
struct alignas(64) PointRec {
  double ere, eim;         // e1 = e^{-i phi} = (cos phi, -sin phi)
  double h0, h1, h2, h3;   // w * Hermite basis at t (dx applied at the fold);
                           // the basis alone (w = 1) for weight-set transforms
  int seg;                 // segment id
  int idx;                 // the point's position in the caller's input arrays
  int pad[2];              // fills the line (written as zero)
};
static_assert(sizeof(PointRec) == 64, "PointRec must be one 64-byte line");

// Per-point record of the spreading scatter, half a line: where the point's
// kernel support starts on the plan's periodic grid of n cells over phi (the
// n u - W/2 of the spreading, u = phi / 2pi, split into an integer cell and
// an FP32 offset in FP64, so the scatter needs no FP64 position) and the
// Hermite weights rounded to FP32.  Written in place of PointRec into the
// same buffer when a transform uses the spreading scatter.
struct alignas(32) SpreadRec {
  float a0, a1, a2, a3;  // w * Hermite basis at t (the basis alone for weight sets)
  int seg;               // segment id
  int idx;               // the point's position in the caller's input arrays
  int j0;                // first cell of the support: ceil(n u - W/2), in [-W/2, n - W/2]
  float z0;              // j0 - n u, in [-W/2, -W/2 + 1)
};
static_assert(sizeof(SpreadRec) == 32, "SpreadRec must be 32 bytes");

}  // namespace dsht
