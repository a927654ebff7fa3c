// Stage (b), part 1: per-point classification, binning and counting sort.
//
// For every point (FP64 throughout, as the binning needs -- t = (x-x_j)/dx
// amplifies the rounding of x by 1/dx):
//   * polar/equatorial split |cos theta| >= xmax          sht/sht.py:128-130
//   * polar points rotated by R_y(+90 deg)               sht/sht.py:134-135
//     (n' = (z, y, -x); any R_z(a) R_y(+-90) gives the same final alm)
//   * theta range checks of basic_sht                    sht/sht.py:171-172
//   * hemisphere split x >= 0 / x < 0                     sht/sht.py:193
//   * bin j = digitize(|x|, x_grid) - 1, t = (|x|-x_j)/dx sht/sht.py:212,221
//   * Hermite weights w * {h0..h3}(t)                     sht/sht.py:239-241
// The points are sorted by segment id (branch, hemisphere, bin) with a
// counting sort whose order within a segment is the input order, so a
// transform's sums -- and its FP32 rounding -- are the same on every call:
//   k_segments:     each point's segment id (the FP64 classification);
//   k_rank_hist:    one wavefront per tile of consecutive input points, in
//                   input order, finds each point's rank among the tile's
//                   points of its segment (same-segment lanes by ballots over
//                   the segment's bits, running counts in the wavefront's own
//                   LDS row: no atomics), and writes the tile's histogram row;
//   k_col_sum, k_scan, k_col_offsets: segment starts and per-(tile, segment)
//                   write offsets;
//   k_place_ranked: every point's record at offset + rank (fully parallel).
// Points whose cos(theta) is NaN are dropped, as the reference's two
// np.nonzero masks drop them (sorted last, never scattered).  Grids with more
// segments than a wavefront's LDS row holds (16-bit counts, Nx > 8192) fall
// back to global atomics (k_hist, k_place), whose order within a segment is
// not fixed.
#include <hip/hip_runtime.h>

#include "common.h"
#include "plan.h"

namespace dsht {

struct PointParams {
  int Nx, nseg, mode;
  int spread_n;  // grid cells of the spreading scatter (SpreadRec records), or 0 (PointRec)
  double xmax, dx, acos_xmax, acos_mxmax;
};

struct Classified {
  int seg;       // nseg = dropped
  int polar;     // 1 if the point went through the rotated branch
  int flag;      // DSHT_STATUS_* bits raised by this point
  double ere, eim, h0, h1, h2, h3;
  int j0;        // spreading support: first cell and offset (when PointParams.spread_n)
  float z0;
};

// kPhase: also the point's azimuth (e^{-i phi} for PointRec, the spreading
// support for SpreadRec); the segment passes need only the segment.
template <bool kPhase = true>
__device__ inline Classified classify(double theta, double phi, double w, const PointParams& p) {
#pragma clang fp contract(off)
  Classified c;
  c.seg = p.nseg;
  c.polar = 0;
  c.flag = 0;
  c.ere = 1.0; c.eim = 0.0;
  c.h0 = c.h1 = c.h2 = c.h3 = 0.0;
  c.j0 = 0;
  c.z0 = 0.f;
  double st, ct;
  double xu, sp, cp;
  int branch;
  if (p.mode == DSHT_MODE_CSHT) {
    ct = fabs(theta);  // the input is cos(theta); one hemisphere by contract
    st = 0.0;
  } else {
    sincos(theta, &st, &ct);
  }
  const double xv = fabs(ct);
  if (p.mode == DSHT_MODE_CSHT) {
    sincos(phi, &sp, &cp);
    xu = ct;
    c.ere = cp;
    c.eim = -sp;
    branch = 0;
  } else if (p.mode == DSHT_MODE_CALL && xv >= p.xmax) {
    // polar: rotate, then basic_sht's checks on the rotated theta
    sincos(phi, &sp, &cp);
    const double X = ct, Y = st * sp, Z = -(st * cp);
    const double rho = sqrt(X * X + Y * Y);
    // cos of the rotated theta, Z / |n| (|n| = 1 up to rounding); the range
    // check theta' in (acos xmax, acos(-xmax)) of basic_sht is the same as
    // -xmax < cos theta' < xmax
    xu = Z / sqrt(rho * rho + Z * Z);
    if (!(xu < p.xmax && xu > -p.xmax)) c.flag |= DSHT_STATUS_POLAR_RANGE;
    if (rho > 0.0) {
      c.ere = X / rho;
      c.eim = -(Y / rho);
    }
    branch = 1;
    c.polar = 1;
  } else if (p.mode == DSHT_MODE_BASIC || xv < p.xmax) {
    if (!(theta >= 0.0 && theta > p.acos_xmax && theta < p.acos_mxmax)) c.flag |= DSHT_STATUS_EQU_RANGE;
    if (kPhase && !p.spread_n) {
      sincos(phi, &sp, &cp);
      c.ere = cp;
      c.eim = -sp;
    }
    xu = ct;
    branch = 0;
  } else {
    return c;  // NaN cos(theta): dropped
  }
  if (xu != xu) return c;  // NaN after rotation: dropped
  if (kPhase && p.spread_n) {
    // u = phi / 2pi mod 1 and the support of the W-tap kernel around n u:
    // for equatorial points from phi itself (FP64, exact to rounding for
    // any moderate phi), for rotated points from the rotated (cos, sin) by
    // atan2
    double u = branch ? atan2(-c.eim, c.ere) * (0.5 / M_PI) : phi * (0.5 / M_PI);
    u -= floor(u);
    if (!(u >= 0.0 && u < 1.0)) u = 0.0;
    const double t = u * (double)p.spread_n;
    const double jf = ceil(t - 0.5 * kSpreadW);
    c.j0 = (int)jf;
    c.z0 = (float)(jf - t);
  }
  const int h = xu < 0.0 ? 1 : 0;
  const double ax = fabs(xu);
  // np.digitize(ax, x_grid) - 1 with x_grid[j] = j * dx (exact FP64 products)
  int j = (int)fmin(ax / p.dx, (double)(p.Nx - 1));
  if (j < 0) j = 0;
  while (j + 1 <= p.Nx - 1 && (double)(j + 1) * p.dx <= ax) ++j;
  while (j > 0 && (double)j * p.dx > ax) --j;
  // |x| in [x_{Nx-1}, xmax): the reference would index past the table; the
  // last interval is used (t slightly above 1).
  if (j > p.Nx - 2) j = p.Nx - 2;
  const double t = (ax - (double)j * p.dx) / p.dx;
  const double omt = 1.0 - t;
  c.h0 = w * ((2.0 * t + 1.0) * (omt * omt));
  c.h1 = w * (t * (omt * omt));
  c.h2 = w * ((t * t) * (3.0 - 2.0 * t));
  c.h3 = w * ((t * t) * (t - 1.0));
  c.seg = (branch * 2 + h) * (p.Nx - 1) + j;
  return c;
}

// The sorted record of a classified point at position pos: a SpreadRec for
// the spreading scatter, else a PointRec.
__device__ inline void place_point(const Classified& c, int sg, int64_t i, int pos,
                                   const PointParams& p, PointRec* __restrict__ rec) {
  if (p.spread_n) {  // a 32-byte SpreadRec, in two 16-byte stores
    float4* d = reinterpret_cast<float4*>(reinterpret_cast<SpreadRec*>(rec) + pos);
    d[0] = make_float4((float)c.h0, (float)c.h1, (float)c.h2, (float)c.h3);
    reinterpret_cast<int4*>(d)[1] = make_int4(sg, (int)i, c.j0, __float_as_int(c.z0));
    return;
  }
  // the whole 64-byte line, in 16-byte stores (a partly written line costs a
  // read-modify-write)
  double2* d = reinterpret_cast<double2*>(rec + pos);
  d[0] = make_double2(c.ere, c.eim);
  d[1] = make_double2(c.h0, c.h1);
  d[2] = make_double2(c.h2, c.h3);
  reinterpret_cast<int4*>(d + 3)[0] = make_int4(sg, (int)i, 0, 0);
}

// Fallback histogram for grids with more segments than k_rank_hist's LDS row
// holds: segment ids (+1 bucket for dropped points) counted with global
// atomics, and the status / polar-count reductions (flags[0] |= status bits,
// flags[1] += polar points).
__global__ __launch_bounds__(256) void k_hist(int64_t Np, const double* __restrict__ theta,
                                              const double* __restrict__ phi, PointParams p,
                                              int* __restrict__ seg_count,
                                              int* __restrict__ flags) {
  int flag = 0, npol = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < Np;
       i += (int64_t)gridDim.x * blockDim.x) {
    Classified c = classify<false>(theta[i], phi[i], 0.0, p);  // weights do not move points
    flag |= c.flag;
    npol += c.polar;
    atomicAdd(&seg_count[c.seg], 1);
  }
  if (flag) atomicOr(&flags[0], flag);
  if (npol) atomicAdd(&flags[1], npol);
}

// Exclusive scans of n ints by one workgroup of 1024 threads: out[i] =
// sum_{k<i} in[k], out[n] = total (out2, if given, gets a copy), and, if
// rank is given, rank[i] = #{k < i : in[k] > 0} for i = 0 .. nrank
// (nrank <= n: the count over the first nrank entries).
// Each thread sums a contiguous slice; the slice totals are scanned within
// each wavefront by shuffles and across the 16 wavefronts in LDS.
__device__ inline int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}

__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ in, int n,
                                               int* __restrict__ out, int* __restrict__ out2,
                                               int* __restrict__ rank, int nrank) {
  __shared__ int wsum[16], wcnt[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int per = (n + 1023) / 1024;
  const int b = min(n, tid * per), e = min(n, b + per);
  int s = 0, c = 0;
  for (int i = b; i < e; ++i) {
    const int v = in[i];
    s += v;
    c += v > 0;
  }
  const int si = wave_incl_scan(s, lane), ci = wave_incl_scan(c, lane);
  if (lane == 63) {
    wsum[w] = si;
    wcnt[w] = ci;
  }
  __syncthreads();
  if (w == 0) {
    const int a = lane < 16 ? wsum[lane] : 0, k = lane < 16 ? wcnt[lane] : 0;
    const int as = wave_incl_scan(a, lane), ks = wave_incl_scan(k, lane);
    if (lane < 16) {
      wsum[lane] = as - a;  // exclusive prefix of the wavefront totals
      wcnt[lane] = ks - k;
    }
  }
  __syncthreads();
  int run = wsum[w] + si - s, cnt = wcnt[w] + ci - c;  // exclusive prefix of this slice
  for (int i = b; i < e; ++i) {
    const int v = in[i];
    out[i] = run;
    if (out2) out2[i] = run;
    if (rank && i <= nrank) rank[i] = cnt;
    run += v;
    cnt += v > 0;
  }
  if (tid == 1023) {
    out[n] = run;
    if (out2) out2[n] = run;
    if (rank && nrank == n) rank[n] = cnt;
  }
}

// max |w| over the valid points of a block into *wmax (float bits of a
// non-negative value: their order as unsigned integers is the float order),
// one atomic per block; non-finite weights are left out (they make the sums
// non-finite by themselves).  The matrix-core scatter scales its fp16
// operands by the power of two this bound gives.
__device__ inline void block_wmax(float v, unsigned* __restrict__ wmax) {
  __shared__ float red[16];
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();  // red[] may still be read by a previous call
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) b = fmaxf(b, red[i]);
    atomic_max_pos(wmax, b);
  }
}

__device__ inline float abs_w(double w, int seg, int nseg) {
  const float a = (float)fabs(w);
  return (seg < nseg && a <= 3.0e38f) ? a : 0.f;
}

// Place each point's 64-byte record at its sorted position.  The order inside
// a segment is that of the atomics (the sums over a segment do not depend on
// it beyond rounding).
__global__ __launch_bounds__(256) void k_place(int64_t Np, const double* __restrict__ theta,
                                               const double* __restrict__ phi,
                                               const double* __restrict__ wt, PointParams p,
                                               int* __restrict__ cursor, PointRec* __restrict__ rec,
                                               unsigned* __restrict__ wmax) {
  float vmax = 0.f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < Np;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double wi = wt ? wt[i] : 1.0;
    Classified c = classify(theta[i], phi[i], wi, p);
    vmax = fmaxf(vmax, abs_w(wi, c.seg, p.nseg));
    const int pos = atomicAdd(&cursor[c.seg], 1);
    place_point(c, c.seg, i, pos, p, rec);
  }
  block_wmax(vmax, wmax);
}

// Segment containing sorted position i: the last s with seg_start[s] <= i
// (empty segments share their start with the next one and are skipped).
__device__ inline int seg_of(const int* seg_start, int nseg, int i) {
  int lo = 0, hi = nseg;  // answer in [0, nseg-1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg_start[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Number of runs (maximal same-segment stretches) in each chunk of the sorted
// valid points: the non-empty segments between the chunk's first and last
// point, found from the segment offsets alone.
__global__ __launch_bounds__(256) void k_count_runs(const int* __restrict__ seg_start,
                                                    const int* __restrict__ ne_rank, int nseg,
                                                    int chunk, int nchunks,
                                                    int* __restrict__ run_count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int nvalid = seg_start[nseg];
  const int b = c * chunk, e = min(nvalid, b + chunk);
  if (b >= e) {
    run_count[c] = 0;
    return;
  }
  const int sb = seg_of(seg_start, nseg, b), se = seg_of(seg_start, nseg, e - 1);
  run_count[c] = ne_rank[se] - ne_rank[sb] + 1;
}

// ---- counting sort with ranks in input order ----
// Tiles of sort_tile(Np) consecutive input points (a multiple of 256), one
// wavefront each, taking 4 x 64 points a step; the next step's coordinates
// are loaded while the current one is ranked.
constexpr int kRankLdsMax = 64 * 1024;  // the wavefront's count row (Nx <= 8192)

// Segment id of every point (input order), and the status / polar-count
// reductions (flags[0] |= status bits, flags[1] += polar points): the FP64
// classification at full occupancy, so the ranking below is integer work.
__global__ __launch_bounds__(256) void k_segments(int64_t Np, const double* __restrict__ theta,
                                                  const double* __restrict__ phi, PointParams p,
                                                  int* __restrict__ seg_of_pt,
                                                  int* __restrict__ flags) {
  __shared__ int red[2][4];
  int flag = 0, npol = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < Np;
       i += (int64_t)gridDim.x * blockDim.x) {
    const Classified c = classify<false>(theta[i], phi[i], 0.0, p);  // weights do not move points
    seg_of_pt[i] = c.seg;
    flag |= c.flag;
    npol += c.polar;
  }
  // one atomic per block (a few thousand blocks): same-address atomics serialize
  for (int off = 32; off > 0; off >>= 1) {
    flag |= __shfl_xor(flag, off);
    npol += __shfl_xor(npol, off);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = flag;
    red[1][threadIdx.x >> 6] = npol;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int f = red[0][0] | red[0][1] | red[0][2] | red[0][3];
    const int n = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    if (f) atomicOr(&flags[0], f);
    if (n) atomicAdd(&flags[1], n);
  }
}

// One wavefront per tile: ranks in input order (4 x 64 points a step, the
// next step's segment ids loaded while the current ones are ranked) and the
// tile's histogram row.
__global__ __launch_bounds__(64) void k_rank_hist(int64_t Np, int tile,
                                                  const int* __restrict__ seg_of_pt, int nb,
                                                  int nbits, int* __restrict__ thist,
                                                  int* __restrict__ rank_of_pt) {
  // counts below 2^16 (a tile holds fewer points): half the LDS of int
  // counters, so twice as many tiles are resident at once
  extern __shared__ unsigned short cnt[];
  const int lane = threadIdx.x;
  for (int i = lane; i < nb; i += 64) cnt[i] = 0;
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * tile, e = min(Np, b + tile);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  // segment ids two steps ahead; the loads are unconditional (index clamped
  // into the tile, validity from the index) so that the compiler can count
  // them and wait for the oldest only
  int sg[4], n1[4], n2[4];
  auto load = [&](int64_t base, int (&v)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = seg_of_pt[min(base + 64 * j + lane, e - 1)];
  };
  load(b, n1);
  load(b + kSortRound, n2);
  for (int64_t base = b; base < e; base += kSortRound) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sg[j] = base + 64 * j + lane < e ? n1[j] : -1;
      n1[j] = n2[j];
    }
    load(base + 2 * kSortRound, n2);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int s = sg[j];
      const bool act = s >= 0;
      // lanes holding the same segment: agree on every bit
      uint64_t peers = __ballot(act);
      for (int bt = 0; bt < nbits; ++bt) {
        const bool bit = (s >> bt) & 1;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      if (act) {
        const int rank = __popcll(peers & below);
        // every lane reads its count before the lowest peer writes the new one
        // (one wavefront: its LDS operations complete in program order)
        const int old = cnt[s];
        if (rank == 0) cnt[s] = (unsigned short)(old + __popcll(peers));
        rank_of_pt[base + 64 * j + lane] = old + rank;
      }
    }
  }
  __syncthreads();
  int* row = thist + (int64_t)blockIdx.x * nb;
  for (int i = lane; i < nb; i += 64) row[i] = cnt[i];
}

// Column passes over thist: 32 segments per block (a half-wavefront reads
// 128 contiguous bytes of a row), each column split over 32 threads (tile
// ranges of ntiles/32 rows, loads unrolled so they are in flight together).
constexpr int kColSeg = 32;
constexpr int kColSplit = 32;

__device__ inline int col_partial(const int* __restrict__ thist, int ntiles, int nb, int s, int q,
                                  int& t0, int& t1) {
  const int per = (ntiles + kColSplit - 1) / kColSplit;
  t0 = min(ntiles, q * per);
  t1 = min(ntiles, t0 + per);
  int acc = 0;
  if (s < nb) {
#pragma unroll 8
    for (int t = t0; t < t1; ++t) acc += thist[(int64_t)t * nb + s];
  }
  return acc;
}

// seg_count[s] = sum over tiles of thist[t][s]
__global__ __launch_bounds__(kColSeg * kColSplit) void k_col_sum(const int* __restrict__ thist,
                                                                 int ntiles, int nb,
                                                                 int* __restrict__ seg_count) {
  __shared__ int part[kColSplit][kColSeg + 1];
  const int sl = threadIdx.x % kColSeg, q = threadIdx.x / kColSeg;
  const int s = blockIdx.x * kColSeg + sl;
  int t0, t1;
  part[q][sl] = col_partial(thist, ntiles, nb, s, q, t0, t1);
  __syncthreads();
  if (q == 0 && s < nb) {
    int tot = 0;
    for (int k = 0; k < kColSplit; ++k) tot += part[k][sl];
    seg_count[s] = tot;
  }
}

// thist[t][s] <- seg_start[s] + sum_{t' < t} thist[t'][s]
__global__ __launch_bounds__(kColSeg * kColSplit) void k_col_offsets(int* __restrict__ thist,
                                                                     int ntiles, int nb,
                                                                     const int* __restrict__ seg_start) {
  __shared__ int part[kColSplit][kColSeg + 1];
  const int sl = threadIdx.x % kColSeg, q = threadIdx.x / kColSeg;
  const int s = blockIdx.x * kColSeg + sl;
  int t0, t1;
  part[q][sl] = col_partial(thist, ntiles, nb, s, q, t0, t1);
  __syncthreads();
  if (s >= nb) return;
  int run = seg_start[s];
  for (int k = 0; k < q; ++k) run += part[k][sl];
  for (int t = t0; t < t1; ++t) {
    const int64_t k = (int64_t)t * nb + s;
    const int c = thist[k];
    thist[k] = run;
    run += c;
  }
}

// Each point's record at its tile's offset for its segment plus its rank.
// Workgroup (tile t, x) takes the points t tile + 1024 x + 256 u + (0 .. 255),
// u < 4: each thread's four points' loads are issued together, so more of
// the scattered record stores overlap the reads and the FP64 classification.
constexpr int kPlacePer = 4;

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup L on XCD
// L % 8) and each XCD has its own L2.  The placement's per-point lookups of
// toff[tile][segment] are random within the tile's row (nseg + 1 ints), so
// the workgroups of one tile are put on one XCD, which then reads the row
// from HBM once for all of them: the 1-D grid is walked in groups of 8 tiles
// x nbx workgroups, XCD = tile % 8.
__device__ inline bool xcd_tile_block(int ntiles, int nbx, int& t, int& bx) {
  const int L = blockIdx.x, per = 8 * nbx;
  const int g = L / per, r = L - g * per;
  t = 8 * g + (r & 7);
  bx = r >> 3;
  return t < ntiles;
}
inline unsigned xcd_tile_grid(int ntiles, int nbx) { return (unsigned)((ntiles + 7) / 8 * 8 * nbx); }


__global__ __launch_bounds__(256) void k_place_ranked(int64_t Np, int tile, int ntiles, int nbx,
                                                      const double* __restrict__ theta,
                                                      const double* __restrict__ phi,
                                                      const double* __restrict__ wt, PointParams p,
                                                      const int* __restrict__ toff,
                                                      const int* __restrict__ seg_of_pt,
                                                      const int* __restrict__ rank_of_pt,
                                                      PointRec* __restrict__ rec,
                                                      unsigned* __restrict__ wmax) {
  const int nb = p.nseg + 1;
  int t, bx;
  if (!xcd_tile_block(ntiles, nbx, t, bx)) return;
  const int64_t b = (int64_t)t * tile, e = min(Np, b + tile);
  const int64_t i0 = b + (int64_t)bx * (256 * kPlacePer) + threadIdx.x;
  double th[kPlacePer], ph[kPlacePer], ww[kPlacePer];
  int sg[kPlacePer], rk[kPlacePer];
#pragma unroll
  for (int u = 0; u < kPlacePer; ++u) {
    const int64_t i = min(i0 + 256 * u, e - 1);  // clamped: loads are unconditional
    th[u] = theta[i];
    ph[u] = phi[i];
    ww[u] = wt ? wt[i] : 1.0;
    sg[u] = seg_of_pt[i];
    rk[u] = rank_of_pt[i];
  }
  // the write offsets, as soon as the segments are known (unconditional:
  // the FP64 classification below runs while they are in flight)
  int off[kPlacePer];
#pragma unroll
  for (int u = 0; u < kPlacePer; ++u) off[u] = toff[(int64_t)t * nb + min(max(sg[u], 0), p.nseg)];
  float vmax = 0.f;
#pragma unroll
  for (int u = 0; u < kPlacePer; ++u) {
    const int64_t i = i0 + 256 * u;
    if (i < e) {
      const Classified c = classify(th[u], ph[u], ww[u], p);
      vmax = fmaxf(vmax, abs_w(ww[u], sg[u], p.nseg));
      place_point(c, sg[u], i, off[u] + rk[u], p, rec);
    }
  }
  block_wmax(vmax, wmax);
}

__global__ __launch_bounds__(256) void k_clear2(int* __restrict__ a, int na, int* __restrict__ b,
                                                int nb, unsigned* __restrict__ c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < na) a[i] = 0;
  if (i < nb) b[i] = 0;
  if (i == 0) c[0] = 0u;
}

// One weight set into the sorted order of the points: wsort[pos] = w[idx].
__global__ __launch_bounds__(256) void k_gather_weights(int64_t Np, const int* __restrict__ nvalid,
                                                        const PointRec* __restrict__ rec,
                                                        const double* __restrict__ w,
                                                        double* __restrict__ wsort) {
  const int64_t pos = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (pos < Np && pos < *nvalid) wsort[pos] = w[rec[pos].idx];
}

int launch_gather_weights(Plan& pl, int64_t Np, const double* w, hipStream_t s) {
  if (Np == 0) return DSHT_OK;
  k_gather_weights<<<(int)((Np + 255) / 256), 256, 0, s>>>(Np, pl.ws.seg_start + pl.nseg, pl.ws.rec,
                                                            w, pl.ws.wsort);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

// ---- the point sets of a dsht_transform_many group, sorted in one pass ----
// Each set's points go to its own range of the record array (global
// positions off[s] ..), sorted by its own segments in input order, exactly as
// launch_points sorts one set; tiles never span sets.  Segment arrays of set
// s start at s (nseg + 2), its flags at 4 s, its max |w| at wmax[s].
__global__ __launch_bounds__(256) void k_clear_many(int n_seg_all, int* __restrict__ seg_count,
                                                    int nflags, int* __restrict__ flags,
                                                    int nsets, unsigned* __restrict__ wmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_seg_all) seg_count[i] = 0;
  if (i < nflags) flags[i] = 0;
  if (i < nsets) wmax[i] = 0u;
}

// segment ids: one block per global tile
__global__ __launch_bounds__(256) void k_segments_many(ManySets ms, PointParams p,
                                                       int* __restrict__ seg_of_pt,
                                                       int* __restrict__ flags) {
  __shared__ int red[2][4];
  const int t = blockIdx.x;
  const int s = many_set_of(ms.tile_off, ms.nsets, t);
  const int lt = t - ms.tile_off[s];
  const int b = lt * ms.tile[s], e = min(ms.n[s], b + ms.tile[s]);
  const double* __restrict__ th = ms.th[s];
  const double* __restrict__ ph = ms.ph[s];
  int flag = 0, npol = 0;
  for (int i = b + threadIdx.x; i < e; i += 256) {
    const Classified c = classify<false>(th[i], ph[i], 0.0, p);
    seg_of_pt[ms.off[s] + i] = c.seg;
    flag |= c.flag;
    npol += c.polar;
  }
  for (int off = 32; off > 0; off >>= 1) {
    flag |= __shfl_xor(flag, off);
    npol += __shfl_xor(npol, off);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = flag;
    red[1][threadIdx.x >> 6] = npol;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int f = red[0][0] | red[0][1] | red[0][2] | red[0][3];
    const int n = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    if (f) atomicOr(&flags[4 * s], f);
    if (n) atomicAdd(&flags[4 * s + 1], n);
  }
}

// ranks in input order, one wavefront per global tile (as k_rank_hist)
__global__ __launch_bounds__(64) void k_rank_hist_many(ManySets ms,
                                                       const int* __restrict__ seg_of_pt, int nb,
                                                       int nbits, int* __restrict__ thist,
                                                       int* __restrict__ rank_of_pt) {
  extern __shared__ unsigned short cnt[];
  const int lane = threadIdx.x;
  for (int i = lane; i < nb; i += 64) cnt[i] = 0;
  __syncthreads();
  const int t = blockIdx.x;
  const int s = many_set_of(ms.tile_off, ms.nsets, t);
  const int lt = t - ms.tile_off[s];
  const int64_t b = ms.off[s] + (int64_t)lt * ms.tile[s];
  const int64_t e = ms.off[s] + min(ms.n[s], (lt + 1) * ms.tile[s]);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int sg[4], n1[4], n2[4];
  auto load = [&](int64_t base, int (&v)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = seg_of_pt[min(base + 64 * j + lane, e - 1)];
  };
  if (b < e) {
    load(b, n1);
    load(b + kSortRound, n2);
  }
  for (int64_t base = b; base < e; base += kSortRound) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sg[j] = base + 64 * j + lane < e ? n1[j] : -1;
      n1[j] = n2[j];
    }
    load(base + 2 * kSortRound, n2);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int sgj = sg[j];
      const bool act = sgj >= 0;
      uint64_t peers = __ballot(act);
      for (int bt = 0; bt < nbits; ++bt) {
        const bool bit = (sgj >> bt) & 1;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      if (act) {
        const int rank = __popcll(peers & below);
        const int old = cnt[sgj];
        if (rank == 0) cnt[sgj] = (unsigned short)(old + __popcll(peers));
        rank_of_pt[base + 64 * j + lane] = old + rank;
      }
    }
  }
  __syncthreads();
  int* row = thist + (int64_t)t * nb;
  for (int i = lane; i < nb; i += 64) row[i] = cnt[i];
}

// column sums / offsets over each set's tiles (grid.y = set)
__global__ __launch_bounds__(kColSeg * kColSplit) void k_col_sum_many(ManySets ms,
                                                                      const int* __restrict__ thist,
                                                                      int nb, int seg_stride,
                                                                      int* __restrict__ seg_count) {
  __shared__ int part[kColSplit][kColSeg + 1];
  const int set = blockIdx.y;
  const int sl = threadIdx.x % kColSeg, q = threadIdx.x / kColSeg;
  const int sg = blockIdx.x * kColSeg + sl;
  const int t0s = ms.tile_off[set], nt = ms.tile_off[set + 1] - t0s;
  int t0, t1;
  part[q][sl] = col_partial(thist + (int64_t)t0s * nb, nt, nb, sg, q, t0, t1);
  __syncthreads();
  if (q == 0 && sg < nb) {
    int tot = 0;
    for (int k = 0; k < kColSplit; ++k) tot += part[k][sl];
    seg_count[set * seg_stride + sg] = tot;
  }
}

__global__ __launch_bounds__(kColSeg * kColSplit) void k_col_offsets_many(
    ManySets ms, int* __restrict__ thist, int nb, int seg_stride,
    const int* __restrict__ seg_start) {
  __shared__ int part[kColSplit][kColSeg + 1];
  const int set = blockIdx.y;
  const int sl = threadIdx.x % kColSeg, q = threadIdx.x / kColSeg;
  const int sg = blockIdx.x * kColSeg + sl;
  const int t0s = ms.tile_off[set], nt = ms.tile_off[set + 1] - t0s;
  int* th = thist + (int64_t)t0s * nb;
  int t0, t1;
  part[q][sl] = col_partial(th, nt, nb, sg, q, t0, t1);
  __syncthreads();
  if (sg >= nb) return;
  int run = seg_start[set * seg_stride + sg];
  for (int k = 0; k < q; ++k) run += part[k][sl];
  for (int t = t0; t < t1; ++t) {
    const int64_t k = (int64_t)t * nb + sg;
    const int c = th[k];
    th[k] = run;
    run += c;
  }
}

// per set (block = set): seg_start = off[s] + exclusive scan of its counts,
// ne_rank = non-empty segments before each
__global__ __launch_bounds__(1024) void k_scan_many(ManySets ms, const int* __restrict__ in,
                                                    int n, int stride, int* __restrict__ out,
                                                    int* __restrict__ rank, int nrank) {
  __shared__ int wsum[16], wcnt[16];
  const int set = blockIdx.x;
  in += set * stride;
  out += set * stride;
  rank += set * stride;
  const int add = ms.off[set];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int per = (n + 1023) / 1024;
  const int b = min(n, tid * per), e = min(n, b + per);
  int sm = 0, c = 0;
  for (int i = b; i < e; ++i) {
    const int v = in[i];
    sm += v;
    c += v > 0;
  }
  const int si = wave_incl_scan(sm, lane), ci = wave_incl_scan(c, lane);
  if (lane == 63) {
    wsum[w] = si;
    wcnt[w] = ci;
  }
  __syncthreads();
  if (w == 0) {
    const int a = lane < 16 ? wsum[lane] : 0, k = lane < 16 ? wcnt[lane] : 0;
    const int as = wave_incl_scan(a, lane), ks = wave_incl_scan(k, lane);
    if (lane < 16) {
      wsum[lane] = as - a;
      wcnt[lane] = ks - k;
    }
  }
  __syncthreads();
  int run = add + wsum[w] + si - sm, cnt = wcnt[w] + ci - c;
  for (int i = b; i < e; ++i) {
    const int v = in[i];
    out[i] = run;
    if (i <= nrank) rank[i] = cnt;
    run += v;
    cnt += v > 0;
  }
  if (tid == 1023) {
    out[n] = run;
    if (nrank == n) rank[n] = cnt;
  }
}

// records at offset + rank: grid (blocks per tile, global tiles)
__global__ __launch_bounds__(256) void k_place_many(ManySets ms, int ntiles, int nbx,
                                                    PointParams p,
                                                    const int* __restrict__ toff,
                                                    const int* __restrict__ seg_of_pt,
                                                    const int* __restrict__ rank_of_pt,
                                                    PointRec* __restrict__ rec,
                                                    unsigned* __restrict__ wmax) {
  const int nb = p.nseg + 1;
  int t, bx;
  if (!xcd_tile_block(ntiles, nbx, t, bx)) return;
  const int s = many_set_of(ms.tile_off, ms.nsets, t);
  const int lt = t - ms.tile_off[s];
  const int b = lt * ms.tile[s], e = min(ms.n[s], b + ms.tile[s]);
  const int i0 = b + bx * (256 * kPlacePer) + threadIdx.x;
  const double* __restrict__ th = ms.th[s];
  const double* __restrict__ ph = ms.ph[s];
  const double* __restrict__ ww = ms.w[s];
  const int o = ms.off[s];
  float vmax = 0.f;
#pragma unroll
  for (int u = 0; u < kPlacePer; ++u) {
    const int i = i0 + 256 * u;
    if (i < e) {
      const double wi = ww[i];
      const int sg = seg_of_pt[o + i], rk = rank_of_pt[o + i];
      const Classified c = classify(th[i], ph[i], wi, p);
      vmax = fmaxf(vmax, abs_w(wi, sg, p.nseg));
      place_point(c, sg, i, toff[(int64_t)t * nb + sg] + rk, p, rec);
    }
  }
  block_wmax(vmax, wmax + s);
}

// runs of each global chunk (chunk c of set s covers global positions
// off[s] + (c - chunk_off[s]) chunk[s] ..)
__global__ __launch_bounds__(256) void k_count_runs_many(ManySets ms,
                                                         const int* __restrict__ seg_start,
                                                         const int* __restrict__ ne_rank, int nseg,
                                                         int stride, int nchunks,
                                                         int* __restrict__ run_count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int s = many_set_of(ms.chunk_off, ms.nsets, c);
  const int* ss = seg_start + s * stride;
  const int* nr = ne_rank + s * stride;
  const int nvalid = ss[nseg];
  const int b = ms.off[s] + (c - ms.chunk_off[s]) * ms.chunk[s];
  const int e = min(nvalid, b + ms.chunk[s]);
  if (b >= e) {
    run_count[c] = 0;
    return;
  }
  const int sb = seg_of(ss, nseg, b), se = seg_of(ss, nseg, e - 1);
  run_count[c] = nr[se] - nr[sb] + 1;
}

int launch_points_many(Plan& pl, const ManySets& ms, int mode, int* flags, unsigned* wmax,
                       hipStream_t s) {
  Workspace& ws = pl.ws;
  PointParams p{pl.Nx, pl.nseg, mode, uses_spread(pl) ? 1 << pl.spread_logn : 0, pl.xmax, pl.dx,
                pl.acos_xmax, pl.acos_mxmax};
  const int nb = pl.nseg + 1, stride = pl.nseg + 2;
  const int ntiles = ms.tile_off[ms.nsets], nchunks = ms.chunk_off[ms.nsets];
  const int npts = ms.off[ms.nsets];
  if (ms.nsets > ws.cap_segsets || ntiles > ws.cap_tiles || nchunks > ws.cap_chunks ||
      npts > ws.cap_points || sizeof(unsigned short) * nb > (size_t)kRankLdsMax) {
    set_error("dsht: point-set group beyond the workspace");
    return DSHT_EINVAL;
  }
  int nbits = 0;
  while ((1 << nbits) < nb) ++nbits;
  const int nclear = std::max(ms.nsets * stride, 4 * ms.nsets);
  k_clear_many<<<(nclear + 255) / 256, 256, 0, s>>>(ms.nsets * stride, ws.seg_count, 4 * ms.nsets,
                                                    flags, ms.nsets, wmax);
  DSHT_LAUNCH_CHECK();
  int* seg_of_pt = reinterpret_cast<int*>(ws.segrank);
  int* rank_of_pt = seg_of_pt + ws.cap_points;
  int maxtile = 0;
  for (int k = 0; k < ms.nsets; ++k) maxtile = std::max(maxtile, ms.tile[k]);
  if (ntiles > 0) {
    k_segments_many<<<ntiles, 256, 0, s>>>(ms, p, seg_of_pt, flags);
    DSHT_LAUNCH_CHECK();
    k_rank_hist_many<<<ntiles, 64, sizeof(unsigned short) * nb, s>>>(ms, seg_of_pt, nb, nbits,
                                                                     ws.thist, rank_of_pt);
    DSHT_LAUNCH_CHECK();
    k_col_sum_many<<<dim3((nb + kColSeg - 1) / kColSeg, ms.nsets), kColSeg * kColSplit, 0, s>>>(
        ms, ws.thist, nb, stride, ws.seg_count);
    DSHT_LAUNCH_CHECK();
  }
  k_scan_many<<<ms.nsets, 1024, 0, s>>>(ms, ws.seg_count, pl.nseg + 1, stride, ws.seg_start,
                                        ws.ne_rank, pl.nseg);
  DSHT_LAUNCH_CHECK();
  if (ntiles > 0) {
    k_col_offsets_many<<<dim3((nb + kColSeg - 1) / kColSeg, ms.nsets), kColSeg * kColSplit, 0, s>>>(
        ms, ws.thist, nb, stride, ws.seg_start);
    DSHT_LAUNCH_CHECK();
    const int nbx = (maxtile + 256 * kPlacePer - 1) / (256 * kPlacePer);
    k_place_many<<<xcd_tile_grid(ntiles, nbx), 256, 0, s>>>(ms, ntiles, nbx, p, ws.thist,
                                                             seg_of_pt, rank_of_pt, ws.rec, wmax);
    DSHT_LAUNCH_CHECK();
  }
  if (nchunks > 0) {
    k_count_runs_many<<<(nchunks + 255) / 256, 256, 0, s>>>(ms, ws.seg_start, ws.ne_rank, pl.nseg,
                                                            stride, nchunks, ws.run_count);
    DSHT_LAUNCH_CHECK();
  }
  k_scan<<<1, 1024, 0, s>>>(ws.run_count, nchunks, ws.run_base, nullptr, nullptr, 0);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

int launch_points(Plan& pl, int64_t Np, const double* theta, const double* phi, const double* wt,
                  int mode, int* flags, hipStream_t s) {
  Workspace& ws = pl.ws;
  PointParams p{pl.Nx, pl.nseg, mode, uses_spread(pl) ? 1 << pl.spread_logn : 0, pl.xmax, pl.dx,
                pl.acos_xmax, pl.acos_mxmax};
  // counters are cleared by a kernel, not a memset node, so the transform can
  // be captured in a HIP graph and replayed
  k_clear2<<<(pl.nseg + 2 + 255) / 256, 256, 0, s>>>(ws.seg_count, pl.nseg + 2, flags, 4, ws.wmax);
  DSHT_LAUNCH_CHECK();
  const int nb = pl.nseg + 1;
  const size_t lds = sizeof(unsigned short) * nb;
  const int tile = sort_tile(Np);
  const int ntiles = (int)((Np + tile - 1) / tile);
  const bool ranked = lds <= (size_t)kRankLdsMax && ntiles <= ws.cap_tiles;
  int nbits = 0;
  while ((1 << nbits) < nb) ++nbits;
  const int threads = 256;
  const int64_t want = (Np + threads - 1) / threads;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, 256 * 16));
  if (Np > 0) {
    if (ranked) {
      int* seg_of_pt = reinterpret_cast<int*>(ws.segrank);
      int* rank_of_pt = seg_of_pt + ws.cap_points;
      k_segments<<<blocks, threads, 0, s>>>(Np, theta, phi, p, seg_of_pt, flags);
      DSHT_LAUNCH_CHECK();
      k_rank_hist<<<ntiles, 64, lds, s>>>(Np, tile, seg_of_pt, nb, nbits, ws.thist, rank_of_pt);
      DSHT_LAUNCH_CHECK();
      k_col_sum<<<(nb + kColSeg - 1) / kColSeg, kColSeg * kColSplit, 0, s>>>(ws.thist, ntiles, nb,
                                                                            ws.seg_count);
    } else {
      k_hist<<<blocks, threads, 0, s>>>(Np, theta, phi, p, ws.seg_count, flags);
    }
    DSHT_LAUNCH_CHECK();
  }
  // segment starts (and cursors) over nseg + 1 buckets (the last, dropped
  // points), and the rank of each non-empty segment among the first nseg
  k_scan<<<1, 1024, 0, s>>>(ws.seg_count, pl.nseg + 1, ws.seg_start, ws.seg_cursor, ws.ne_rank,
                            pl.nseg);
  DSHT_LAUNCH_CHECK();
  if (Np > 0) {
    if (ranked) {
      k_col_offsets<<<(nb + kColSeg - 1) / kColSeg, kColSeg * kColSplit, 0, s>>>(
          ws.thist, ntiles, nb, ws.seg_start);
      DSHT_LAUNCH_CHECK();
      const int* seg_of_pt = reinterpret_cast<const int*>(ws.segrank);
      const int nbx = (tile + 256 * kPlacePer - 1) / (256 * kPlacePer);
      k_place_ranked<<<xcd_tile_grid(ntiles, nbx), 256, 0, s>>>(
          Np, tile, ntiles, nbx, theta, phi, wt, p, ws.thist, seg_of_pt, seg_of_pt + ws.cap_points,
          ws.rec, ws.wmax);
    } else {
      k_place<<<blocks, threads, 0, s>>>(Np, theta, phi, wt, p, ws.seg_cursor, ws.rec, ws.wmax);
    }
    DSHT_LAUNCH_CHECK();
  }
  const int nchunks = (int)((Np + pl.chunk - 1) / pl.chunk);
  if (nchunks > 0) {
    k_count_runs<<<(nchunks + 255) / 256, 256, 0, s>>>(ws.seg_start, ws.ne_rank, pl.nseg, pl.chunk,
                                                      nchunks, ws.run_count);
    DSHT_LAUNCH_CHECK();
  }
  k_scan<<<1, 1024, 0, s>>>(ws.run_count, nchunks, ws.run_base, nullptr, nullptr, 0);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
