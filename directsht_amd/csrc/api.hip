// C ABI of libdsht.so: plan management, transform orchestration and the
// reference-compatible make_table / do_transform entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "common.h"
#include "plan.h"

namespace dsht {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

void build_layout(int Nl, int Npair, std::vector<GroupDesc>& groups, std::vector<int64_t>& mp_off,
                  int64_t& tab_vec4) {
  groups.clear();
  mp_off.assign(2 * (size_t)Nl, 0);
  int64_t off = 0;
  for (int m = 0; m < Nl; ++m) {
    for (int P = 0; P < 2; ++P) {
      mp_off[2 * m + P] = off;
      const int R = rows_mp(Nl, m, P);
      for (int i0 = 0; i0 < R; i0 += kGroupRows) {
        GroupDesc g;
        g.m = m;
        g.P = P;
        g.i0 = i0;
        g.nrows = std::min(kGroupRows, R - i0);
        g.off = off;
        groups.push_back(g);
        off += (int64_t)g.nrows * Npair;
      }
    }
  }
  tab_vec4 = off;
}

// The matrix-core contraction's work items: each group of 64 rows together
// with the next group of the same (m, P) (which starts nrows * Npair table
// entries further on), else alone.
static std::vector<GroupDesc> pair_groups(const std::vector<GroupDesc>& groups) {
  std::vector<GroupDesc> out;
  for (size_t i = 0; i < groups.size(); ++i) {
    GroupDesc g = groups[i];
    if (i + 1 < groups.size() && g.nrows == kGroupRows && groups[i + 1].m == g.m &&
        groups[i + 1].P == g.P) {
      g.nrows += groups[++i].nrows;
    }
    out.push_back(g);
  }
  return out;
}

// Launch order of the contraction's row groups.  Workgroups are dispatched to
// the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and each XCD has its
// own L2.  Give every XCD a contiguous, row-balanced range of groups, so the
// groups of one (m, P) -- which all read the same node coefficients -- run
// on one XCD and share them in its L2; pad the shorter ranges with empty
// groups (nrows = 0), which exit at once.
std::vector<GroupDesc> xcd_schedule(const std::vector<GroupDesc>& groups, int nxcd) {
  int64_t total = 0;
  for (const GroupDesc& g : groups) total += g.nrows;
  std::vector<std::vector<GroupDesc>> per(nxcd);
  int64_t before = 0;
  for (const GroupDesc& g : groups) {
    const int64_t mid2 = 2 * before + g.nrows;  // twice the group's row midpoint
    int x = (int)(mid2 * nxcd / std::max<int64_t>(1, 2 * total));
    x = std::min(nxcd - 1, std::max(0, x));
    per[x].push_back(g);
    before += g.nrows;
  }
  size_t len = 0;
  for (const auto& v : per) len = std::max(len, v.size());
  GroupDesc empty{};
  std::vector<GroupDesc> out(len * nxcd, empty);
  for (int x = 0; x < nxcd; ++x)
    for (size_t q = 0; q < per[x].size(); ++q) out[q * nxcd + x] = per[x][q];
  return out;
}

static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

struct WsSizes {
  int64_t points, chunks, runs, tiles;
};

static WsSizes ws_sizes(int Nx, int64_t Np, int chunk) {
  WsSizes z;
  z.points = std::max<int64_t>(Np, 1);
  // the most work items of any transform up to Np points (work_chunk)
  z.chunks = std::max<int64_t>((Np + chunk - 1) / chunk,
                               std::min<int64_t>((Np + kChunkMin - 1) / kChunkMin, kChunkTarget));
  z.runs = z.chunks + num_segments(Nx) + 1;
  // tiles of the largest transform up to Np points
  z.tiles = sort_tiles_cap(Np);
  return z;
}

int64_t workspace_bytes(int Nl, int Nx, int precision, int64_t Np, int chunk) {
  const WsSizes z = ws_sizes(Nx, Np, chunk);
  const int Npair = node_pairs(Nx);
  const int64_t Nlm = (int64_t)Nl * (Nl + 1) / 2;
  const int nseg = num_segments(Nx);
  int64_t b = 0;
  b += align256(z.points * sizeof(PointRec));
  b += 2 * align256((z.chunks + 1) * sizeof(int));
  b += align256(z.runs * sizeof(int));
  b += align256(z.runs * (int64_t)Nl * 8 * (precision == DSHT_FP32 ? 4 : 8));
  b += 6 * align256((nseg + 2) * sizeof(int));
  b += align256(z.tiles * (nseg + 1) * sizeof(int));
  b += align256(z.points * (int64_t)sizeof(int2));
  b += align256((int64_t)Nl * 2 * Npair * 24 * (precision == DSHT_FP32 ? 4 : 8));
  b += 2 * align256(Nlm * 2 * sizeof(double));
  b += align256((int64_t)kContractSplits * Nlm * 8 * sizeof(double));
  b += 256;
  return b;
}

static void free_workspace(Workspace& ws) {
  if (ws.fixed) {  // carved from the fixed buffer: nothing to free
    ws = Workspace{};
    return;
  }
  for (void* p : {(void*)ws.rec, (void*)ws.run_count, (void*)ws.run_base, (void*)ws.run_seg,
                  ws.part, (void*)ws.seg_count, (void*)ws.seg_start, (void*)ws.seg_cursor,
                  (void*)ws.seg_run_lo, (void*)ws.seg_run_hi,
                  (void*)ws.ne_rank, (void*)ws.thist, (void*)ws.segrank, ws.rhs, (void*)ws.rmax, (void*)ws.wmax, (void*)ws.wsort,
                  (void*)ws.sidx, (void*)ws.wsf,
                  (void*)ws.a_equ, (void*)ws.a_pol, (void*)ws.a_tmp, (void*)ws.a_der, (void*)ws.flags})
    if (p) (void)hipFree(p);
  ws = Workspace{};
}

#define DSHT_ALLOC(ptr, nbytes_)                                                        \
  do {                                                                                  \
    const size_t n_ = (size_t)(nbytes_);                                                \
    hipError_t e_ = hipMalloc((void**)&(ptr), std::max<size_t>(n_, 16));                \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string("hipMalloc(") + std::to_string(n_) + " bytes): " +         \
                hipGetErrorString(e_));                                                 \
      return DSHT_ENOMEM;                                                               \
    }                                                                                   \
    ws.bytes += (int64_t)n_;                                                            \
  } while (0)

// Grow the workspace to hold a transform of Np points (per set) with nsets
// sets, and the sorted-weight array if `wsort` is asked for.  Growth waits
// for the device so buffers still in use by queued work are never freed
// under it.
// Node-coefficient elements per (m, P, node pair) for a transform of nsets
// sets: one 24-value slot for the plain contraction, 16 per set for the
// matrix-core contraction.
static int64_t rhs_per_pair(const Plan& pl, int nsets) {
  return use_mfma_contract(pl, nsets) ? std::max<int64_t>(24, 16 * (int64_t)nsets) : 24;
}

// Lay out every workspace array of a transform with capacities `want`
// through alloc(ptr, bytes) (device allocation, or carving from a fixed
// buffer), and record the capacities.  psets > 1: weight-set slots of
// slot.points points and slot.runs runs each; sorted weight rows (wsf) for
// max(psets, wrows) sets of slot.points points if that is more than 1.
template <class Alloc>
static int layout_workspace(const Plan& pl, Workspace& ws, const WsSizes& want, int sets,
                            int psets, bool with_w, int64_t rhs_elems, Alloc&& alloc,
                            int segsets = 1, WsSizes slot = WsSizes{0, 0, 0, 0}, int wrows = 0) {
  const int wr = std::max(psets > 1 ? psets : 0, wrows);
  if (psets <= 1) slot.runs = 0;
  if (wr <= 1) slot = WsSizes{0, 0, 0, 0};
  const int fsz = pl.precision == DSHT_FP32 ? 4 : 8;
  int rc = DSHT_OK;
  auto get = [&](auto*& ptr, int64_t nbytes) {
    void* q = nullptr;
    if (rc == DSHT_OK) rc = alloc(&q, nbytes);
    ptr = static_cast<std::remove_reference_t<decltype(ptr)>>(q);
  };
  get(ws.rec, want.points * (int64_t)sizeof(PointRec));
  get(ws.run_count, (want.chunks + 1) * (int64_t)sizeof(int));
  get(ws.run_base, (want.chunks + 1) * (int64_t)sizeof(int));
  get(ws.run_seg, want.runs * (int64_t)sizeof(int));
  get(ws.part, std::max(want.runs, psets * slot.runs) * (int64_t)pl.Nl * 8 * fsz);
  get(ws.seg_count, (int64_t)segsets * (pl.nseg + 2) * (int64_t)sizeof(int));
  get(ws.seg_start, (int64_t)segsets * (pl.nseg + 2) * (int64_t)sizeof(int));
  get(ws.seg_cursor, (int64_t)segsets * (pl.nseg + 2) * (int64_t)sizeof(int));
  get(ws.seg_run_lo, (int64_t)segsets * (pl.nseg + 2) * (int64_t)sizeof(int));
  get(ws.seg_run_hi, (int64_t)segsets * (pl.nseg + 2) * (int64_t)sizeof(int));
  get(ws.ne_rank, (int64_t)segsets * (pl.nseg + 2) * (int64_t)sizeof(int));
  get(ws.thist, want.tiles * (pl.nseg + 1) * (int64_t)sizeof(int));
  get(ws.segrank, want.points * (int64_t)sizeof(int2));
  get(ws.rhs, rhs_elems * fsz);
  get(ws.rmax, sets * (int64_t)sizeof(unsigned));
  get(ws.wmax, sets * (int64_t)sizeof(unsigned));
  if (wr > 1) {
    ws.ldws = (slot.points + 63) / 64 * 64 + 64;
    get(ws.sidx, want.points * (int64_t)sizeof(int));
    // sorted rows [wr][ldws], then the point-major copy [slot.points][wr]
    get(ws.wsf, 2 * wr * ws.ldws * (int64_t)sizeof(float));
  }
  if (with_w) get(ws.wsort, want.points * (int64_t)sizeof(double));
  get(ws.a_equ, pl.Nlm * 2 * sets * (int64_t)sizeof(double));
  get(ws.a_pol, pl.Nlm * 2 * sets * (int64_t)sizeof(double));
  // the batched rotation's input, in groups of up to 16 sets
  if (sets > 1) get(ws.a_tmp, pl.Nlm * 2 * ((sets + 15) / 16 * 16) * (int64_t)sizeof(double));
  get(ws.a_der, (int64_t)kContractSplits * pl.Nlm * 8 * (int64_t)sizeof(double));
  get(ws.flags, 4 * sets * (int64_t)sizeof(int));
  if (rc) return rc;
  ws.cap_points = want.points;
  ws.cap_chunks = want.chunks;
  ws.cap_runs = want.runs;
  ws.cap_tiles = want.tiles;
  ws.cap_sets = sets;
  ws.cap_part_sets = psets;
  ws.cap_wsf_sets = wr > 1 ? wr : 0;
  ws.cap_set_points = slot.points;
  ws.cap_set_runs = slot.runs;
  ws.has_wsort = with_w;
  ws.rhs_elems = rhs_elems;
  ws.cap_segsets = segsets;
  return DSHT_OK;
}

// Grow the workspace to capacities z (and the other counts) at least;
// part_sets > 1: as many weight-set slots, each for a transform within z;
// wrows > 1: as many sorted weight rows of z.points points.
static int ensure_workspace_z(Plan& pl, const WsSizes& z, int nsets, bool wsort, int part_sets,
                              int segsets, int wrows = 0) {
  Workspace& ws = pl.ws;
  const int64_t rhs_need = (int64_t)pl.Nl * 2 * pl.Npair * rhs_per_pair(pl, nsets);
  const bool slots_ok =
      (part_sets <= 1 || (z.points <= ws.cap_set_points && z.runs <= ws.cap_set_runs)) &&
      (wrows <= 1 || (z.points <= ws.cap_set_points && wrows <= ws.cap_wsf_sets));
  if (ws.flags && z.points <= ws.cap_points && z.chunks <= ws.cap_chunks && z.runs <= ws.cap_runs &&
      z.tiles <= ws.cap_tiles && nsets <= ws.cap_sets && (!wsort || ws.has_wsort) &&
      rhs_need <= ws.rhs_elems && part_sets <= ws.cap_part_sets && segsets <= ws.cap_segsets &&
      slots_ok)
    return DSHT_OK;
  if (ws.fixed) {
    set_error("dsht: transform larger than the fixed workspace it was given");
    return DSHT_EINVAL;
  }
  DSHT_HIP(hipDeviceSynchronize());
  WsSizes want{std::max(z.points, ws.cap_points), std::max(z.chunks, ws.cap_chunks),
               std::max(z.runs, ws.cap_runs), std::max(z.tiles, ws.cap_tiles)};
  const int sets = std::max(nsets, ws.cap_sets);
  const int psets = std::max(part_sets, ws.cap_part_sets);
  const bool with_w = wsort || ws.has_wsort;
  const int64_t rhs_elems = std::max(rhs_need, ws.rhs_elems);
  const int ssets = std::max(segsets, ws.cap_segsets);
  WsSizes slot{ws.cap_set_points, 0, ws.cap_set_runs, 0};
  if (part_sets > 1 || wrows > 1) slot.points = std::max(slot.points, z.points);
  if (part_sets > 1) slot.runs = std::max(slot.runs, z.runs);
  const int wr = std::max(wrows, ws.cap_wsf_sets);
  free_workspace(ws);
  auto dev_alloc = [&](void** q, int64_t nbytes) -> int {
    const size_t n = (size_t)std::max<int64_t>(nbytes, 16);
    hipError_t e = hipMalloc(q, n);
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc(") + std::to_string(nbytes) + " bytes): " +
                hipGetErrorString(e));
      return DSHT_ENOMEM;
    }
    ws.bytes += nbytes;
    return DSHT_OK;
  };
  return layout_workspace(pl, ws, want, sets, psets, with_w, rhs_elems, dev_alloc, ssets, slot,
                          wr);
}

static int ensure_workspace(Plan& pl, int64_t Np, int nsets = 1, bool wsort = false,
                            int part_sets = 1, int wrows = 0) {
  return ensure_workspace_z(pl, ws_sizes(pl.Nx, Np, pl.chunk_max), nsets, wsort, part_sets, 1,
                            wrows);
}

// Point-set batches (dsht_transform_many) on the matrix-core engines run in
// groups of consecutive sets sorted, scattered and folded together: at most
// kManyGroup sets holding at most kManyPoints points (a larger set forms a
// group of its own), and few enough sets that their run partials -- besides
// the work items, up to nseg + 2 runs per set -- stay within kManyPartBytes.
constexpr int64_t kManyPoints = (int64_t)1 << 23;
constexpr int64_t kManyPartBytes = (int64_t)4 << 30;
constexpr size_t kManyLdsMax = 64 * 1024;  // the rank pass's 16-bit count row

static int many_group_cap(const Plan& pl) {
  const int64_t per_set = (int64_t)(pl.nseg + 2) * pl.Nl * 8 * sizeof(float);
  return (int)std::max<int64_t>(1, std::min<int64_t>(kManyGroup, kManyPartBytes / per_set));
}

static bool uses_many(const Plan& pl, int nsets, int mode) {
  return use_mfma_contract(pl, nsets) &&
         (pl.scatter == DSHT_SCATTER_MATRIX || pl.scatter == DSHT_SCATTER_AUTO) &&
         mode != DSHT_MODE_CSHT && sizeof(unsigned short) * (pl.nseg + 1) <= kManyLdsMax &&
         many_group_cap(pl) > 1;
}

// capacities that hold every group of any batch of nsets sets of up to Np
// points each: a group holds at most T points, its work items are of
// work_chunk(chunk_max, group points) points and its tiles of
// sort_tile(group points), each set adding at most one partial item and
// tile, and a set of n points has at most min(n, nseg + 2) runs beyond its
// work items
static WsSizes many_sizes(const Plan& pl, int64_t Np, int nsets) {
  const int g = std::min(nsets, many_group_cap(pl));
  const int64_t T = std::max<int64_t>(Np, std::min<int64_t>((int64_t)g * Np, kManyPoints));
  const WsSizes one = ws_sizes(pl.Nx, T, pl.chunk_max);
  WsSizes z;
  z.points = one.points;
  z.tiles = sort_tiles_cap(T) + g;
  z.chunks = one.chunks + g;
  z.runs = z.chunks + std::min<int64_t>(T, (int64_t)g * (pl.nseg + 2));
  return z;
}

// A workspace for one single-set transform of Np points carved from the
// fixed device buffer [base, base + cap): the same arrays as ensure_workspace
// lays out, each 256-byte aligned.  Nothing is allocated.
// *need receives the bytes the layout takes (also when they exceed cap).
static int carve_workspace(const Plan& pl, int64_t Np, char* base, int64_t cap, Workspace& ws,
                           int64_t* need) {
  const WsSizes z = ws_sizes(pl.Nx, Np, pl.chunk_max);
  const int64_t rhs_need = (int64_t)pl.Nl * 2 * pl.Npair * rhs_per_pair(pl, 1);
  int64_t off = 0;
  auto bump = [&](void** q, int64_t nbytes) -> int {
    const int64_t n = align256(std::max<int64_t>(nbytes, 16));
    *q = off + n <= cap ? static_cast<void*>(base + off) : nullptr;
    off += n;
    return DSHT_OK;
  };
  ws = Workspace{};
  layout_workspace(pl, ws, z, 1, 1, false, rhs_need, bump);
  ws.fixed = true;
  if (need) *need = off;
  return off <= cap ? DSHT_OK : DSHT_ENOMEM;
}

// ---------------------------------------------------------------------------
// Fixed scratch buffer of dsht_transform_noalloc: kScratchBytes of device
// memory in the library's code object, held on every device the library is
// loaded on until the process ends.  Host code never passes the symbol
// itself to a kernel: its device address, one per device, comes from
// hipGetSymbolAddress on that device's first call.
// ---------------------------------------------------------------------------
constexpr int64_t kScratchBytes = (int64_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_dev[kMaxDevices] = {};  // per device: address of `scratch`

// The calling thread's current device must be `dev`.
static int scratch_address(int dev, char** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("dsht_transform_noalloc: device index beyond the supported range");
    return DSHT_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("dsht_transform_noalloc: hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(e));
      return DSHT_EHIP;
    }
    g_scratch_dev[dev] = static_cast<char*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return DSHT_OK;
}

// d^l_{l,m}(pi/2) = (-1)^{l-m} sqrt(C(2l, l+m)) / 2^l as mantissa * 2^exp, and
// the recursion coefficients A(m')/B(m') and 1/B(m'), for every l < Nl.
static void rotation_constants(int Nl, std::vector<double>& top, std::vector<int>& ex,
                               std::vector<double>& A, std::vector<double>& iB) {
  const int64_t n = (int64_t)Nl * (Nl + 1) / 2;
  top.assign(n, 0.0);
  ex.assign(n, 0);
  A.assign(n, 0.0);
  iB.assign(n, 0.0);
  double d0 = 1.0;  // d^l_{l,0}
  for (int l = 0; l < Nl; ++l) {
    if (l > 0) d0 *= -std::sqrt((2.0 * l - 1.0) / (2.0 * l));
    const int64_t tri = (int64_t)l * (l + 1) / 2;
    double v = d0;
    int e = 0;
    for (int m = 0; m <= l; ++m) {
      top[tri + m] = v;
      ex[tri + m] = e;
      if (m < l) {
        v *= -std::sqrt((double)(l - m) / (double)(l + m + 1));
        if (std::fabs(v) < 0x1p-400) {
          v *= 0x1p400;
          e -= 400;
        }
      }
    }
    for (int mp = 1; mp <= l; ++mp) {
      const double b = std::sqrt((double)(l + mp) * (double)(l - mp + 1));
      A[tri + mp] = std::sqrt((double)(l - mp) * (double)(l + mp + 1)) / b;
      iB[tri + mp] = 1.0 / b;
    }
  }
}

// Coefficients of the normalized Legendre recurrence the FP32 contraction
// runs (k_contract_otf), per order m at otf_offset(m, Nl), pairs {A, B}:
//   k = 0:  A = (-1)^m K_m N_m,  Y_mm(x) = A (1 - x^2)^(m/2),
//           K_m = prod_{j<=m} sqrt(1 - 1/(2j))          (sht/legendre_py.py:28-30)
//   k = 1:  A = sqrt(2m+1) N_{m+1} / N_m,  Y_{m+1,m}(x) = A x Y_mm(x)   (:32-34)
//   k >= 2, l = m + k < Nl:  Y_l = A x Y_{l-1} - B Y_{l-2} with
//           A = f1 (2l-1) / (l-m) N_l / N_{l-1},
//           B = f1 f2 (l+m-1) / (l-m) N_l / N_{l-2}                 (:91-100)
//   f1 = sqrt((l-m)/(l+m)), f2 = sqrt((l-m-1)/(l+m-1)), N_l = sqrt((2l+1)/4pi)
//   (sht/legendre_py.py:77-81); zero for Nl <= l < Nl + 30 (rows past the band
//   limit, which the kernel generates in its last step and discards).
static void otf_coefficients(int Nl, std::vector<double>& rc) {
  rc.assign(2 * (size_t)otf_offset(Nl, Nl), 0.0);
  auto N = [](int l) { return std::sqrt((double)(2 * l + 1) / 4.0 / M_PI); };
  double K = 1.0;
  for (int m = 0; m < Nl; ++m) {
    if (m > 0) K *= std::sqrt(1.0 - 1.0 / (double)(2 * m));
    double* r = rc.data() + 2 * otf_offset(m, Nl);
    r[0] = (m & 1 ? -K : K) * N(m);
    r[2] = std::sqrt(2.0 * m + 1.0) * N(m + 1) / N(m);
    for (int l = m + 2; l < Nl; ++l) {
      const double f1 = std::sqrt((double)(l - m) / (double)(l + m));
      const double f2 = std::sqrt(((double)(l - m) - 1.0) / ((double)(l + m) - 1.0));
      r[2 * (l - m)] = f1 * (double)(2 * l - 1) / (double)(l - m) * N(l) / N(l - 1);
      r[2 * (l - m) + 1] = f1 * f2 * (double)(l + m - 1) / (double)(l - m) * N(l) / N(l - 2);
    }
  }
}

// Spreading-scatter constants: n = 2^logn >= 4 Nl (at least 256), beta =
// 2.3 W, the twiddles e^{-2 pi i k / n} and 1/psihat(m) for m < Nl, with
// psihat(m) = (W/2) int_{-1}^{1} exp(beta (sqrt(1 - z^2) - 1)) cos(pi m W z / n) dz
// by 128-point Gauss-Legendre quadrature (FP64).
static int spread_constants(int Nl, int& logn, double& beta, std::vector<float2>& tw,
                            std::vector<float>& ipsi) {
  logn = 8;
  while ((1 << logn) < 4 * Nl) ++logn;
  const int n = 1 << logn;
  const double W = (double)kSpreadW;
  beta = 2.3 * W;
  tw.resize(n);
  for (int k = 0; k < n; ++k) {
    const double a = 2.0 * M_PI * (double)k / (double)n;
    tw[k] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  constexpr int kq = 128;
  double zq[kq], wq[kq];
  for (int i = 0; i < kq; ++i) {  // Legendre roots by Newton's method
    double z = std::cos(M_PI * (i + 0.75) / (kq + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= kq; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = kq * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    zq[i] = z;
    wq[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  // the transform of the kernel the device evaluates (its sqrt and exp are
  // the hardware's, not correctly rounded): its values at the nodes, taken
  // at the nodes rounded to FP32 as the scatters' offsets are
  float zf[kq], psi[kq];
  for (int i = 0; i < kq; ++i) zf[i] = (float)zq[i];
  if (int rc = spread_kernel_values(zf, kq, beta, psi)) return rc;
  ipsi.resize(Nl);
  for (int m = 0; m < Nl; ++m) {
    double sum = 0.0;
    for (int i = 0; i < kq; ++i) sum += wq[i] * (double)psi[i] * std::cos(M_PI * m * W * zf[i] / n);
    ipsi[m] = (float)(1.0 / (0.5 * W * sum));
  }
  return DSHT_OK;
}

// alloc_tables: the plan's Y-only table (FP64 or FP32 by precision), built
// if `build`; caller: instead an FP64 {Y, Y'} table that the caller's tables
// are copied into (csht do_transform).
static int plan_init(Plan& pl, int Nl, int Nx, double xmax, int precision, bool alloc_tables,
                     bool build, bool caller = false) {
  if (Nl < 2 || Nx < 2 || !(xmax > 0.0 && xmax < 1.0)) {
    set_error("dsht: need Nl >= 2, Nx >= 2 and 0 < xmax < 1");
    return DSHT_EINVAL;
  }
  if (Nx > kMaxNx) {
    set_error("dsht: Nx > " + std::to_string(kMaxNx) + " is not supported");
    return DSHT_EINVAL;
  }
  // tables alone (make_table) have no band limit beyond memory; transforms
  // are bounded by the polar rotation's per-degree LDS rows
  if (alloc_tables && Nl > kMaxNl) {
    set_error("dsht: transforms need Nl <= " + std::to_string(kMaxNl));
    return DSHT_EINVAL;
  }
  if (precision != DSHT_FP64 && precision != DSHT_FP32) {
    set_error("dsht: precision must be DSHT_FP64 or DSHT_FP32");
    return DSHT_EINVAL;
  }
  DSHT_HIP(hipGetDevice(&pl.device));
  pl.Nl = Nl;
  pl.Nx = Nx;
  pl.Npair = node_pairs(Nx);
  pl.nseg = num_segments(Nx);
  pl.Nlm = (int64_t)Nl * (Nl + 1) / 2;
  pl.xmax = xmax;
  pl.dx = xmax / (double)(Nx - 1);
  pl.acos_xmax = std::acos(xmax);
  pl.acos_mxmax = std::acos(-xmax);
  pl.precision = precision;
  std::vector<GroupDesc> groups;
  std::vector<int64_t> mp_off;
  build_layout(Nl, pl.Npair, groups, mp_off, pl.tab_vec4);
  std::vector<GroupDesc> mgroups = xcd_schedule(pair_groups(groups), kXcds);
  groups = xcd_schedule(groups, kXcds);
  pl.ngroups = (int)groups.size();
  pl.nmgroups = (int)mgroups.size();
  groups.insert(groups.end(), mgroups.begin(), mgroups.end());
  Workspace& ws = pl.ws;  // DSHT_ALLOC counts bytes into ws.bytes; moved below
  DSHT_ALLOC(pl.groups, groups.size() * sizeof(GroupDesc));
  pl.mgroups = pl.groups + pl.ngroups;
  DSHT_ALLOC(pl.mp_off, mp_off.size() * sizeof(int64_t));
  DSHT_HIP(hipMemcpy(pl.groups, groups.data(), groups.size() * sizeof(GroupDesc), hipMemcpyHostToDevice));
  DSHT_HIP(hipMemcpy(pl.mp_off, mp_off.data(), mp_off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  pl.caller_tables = caller;
  if (caller) DSHT_ALLOC(pl.tab64, pl.tab_vec4 * 4 * sizeof(double));
  if (alloc_tables && !caller && precision == DSHT_FP64)
    DSHT_ALLOC(pl.tab64y, pl.tab_vec4 * 2 * sizeof(double));
  if (alloc_tables && !caller && precision == DSHT_FP32) {
    // no Y table: the contraction regenerates Y from the recurrence's coefficients
    std::vector<double> rcoef;
    otf_coefficients(Nl, rcoef);
    pl.otf = true;
    DSHT_ALLOC(pl.otf_rc, rcoef.size() * sizeof(double));
    DSHT_HIP(hipMemcpy(pl.otf_rc, rcoef.data(), rcoef.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  std::vector<double> top, A, iB;
  std::vector<int> ex;
  rotation_constants(Nl, top, ex, A, iB);
  DSHT_ALLOC(pl.rot_top, top.size() * sizeof(double));
  DSHT_ALLOC(pl.rot_exp, ex.size() * sizeof(int));
  DSHT_ALLOC(pl.rot_AB, A.size() * sizeof(double));
  DSHT_ALLOC(pl.rot_iB, iB.size() * sizeof(double));
  DSHT_HIP(hipMemcpy(pl.rot_top, top.data(), top.size() * sizeof(double), hipMemcpyHostToDevice));
  DSHT_HIP(hipMemcpy(pl.rot_exp, ex.data(), ex.size() * sizeof(int), hipMemcpyHostToDevice));
  DSHT_HIP(hipMemcpy(pl.rot_AB, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice));
  DSHT_HIP(hipMemcpy(pl.rot_iB, iB.data(), iB.size() * sizeof(double), hipMemcpyHostToDevice));
  if (precision == DSHT_FP32) {
    std::vector<float2> tw;
    std::vector<float> ipsi;
    if (int rc = spread_constants(Nl, pl.spread_logn, pl.spread_beta, tw, ipsi)) return rc;
    DSHT_ALLOC(pl.spread_tw, tw.size() * sizeof(float2));
    DSHT_ALLOC(pl.spread_ipsi, ipsi.size() * sizeof(float));
    DSHT_HIP(hipMemcpy(pl.spread_tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    DSHT_HIP(hipMemcpy(pl.spread_ipsi, ipsi.data(), ipsi.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  pl.table_bytes = ws.bytes;
  ws.bytes = 0;
  pl.table_hold = std::make_shared<TableOwner>(
      pl.device, std::vector<void*>{pl.groups, pl.mp_off, pl.tab64y, pl.tab32y, pl.otf_rc, pl.tab64,
                                    pl.rot_top, pl.rot_exp, pl.rot_AB, pl.rot_iB, pl.spread_tw,
                                    pl.spread_ipsi});
  if (alloc_tables && build && !caller) {
    int rc = precision == DSHT_FP64 ? launch_tables(pl, 0, pl.tab64y, nullptr, 0) : DSHT_OK;
    if (rc) return rc;
    DSHT_HIP(hipDeviceSynchronize());
  }
  return DSHT_OK;
}

// The FP32 {Y, Y'} table of the matrix-core contraction: built on first use
// (shared by the plan's clones); waits for the device once.
static int ensure_full32(Plan& pl, const float** out) {
  TableOwner& o = *pl.table_hold;
  std::lock_guard<std::mutex> lock(o.mu);
  if (!o.full32) {
    const int64_t bytes = pl.tab_vec4 * 4 * (int64_t)sizeof(float);
    float* t = nullptr;
    hipError_t e = hipMalloc((void**)&t, (size_t)bytes);
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc(") + std::to_string(bytes) + " bytes) for the {Y, Y'} table: " +
                hipGetErrorString(e));
      return DSHT_ENOMEM;
    }
    int rc = launch_tables(pl, 2, t, nullptr, 0);
    if (rc == DSHT_OK && hipDeviceSynchronize() != hipSuccess) {
      set_error("dsht: building the {Y, Y'} table failed");
      rc = DSHT_EHIP;
    }
    if (rc) {
      (void)hipFree(t);
      return rc;
    }
    o.full32 = t;
    o.full32_bytes = bytes;
  }
  *out = o.full32;
  return DSHT_OK;
}

static void plan_free(Plan& pl) {
  free_workspace(pl.ws);
  for (hipEvent_t e : pl.events) (void)hipEventDestroy(e);
  if (pl.table_hold) {
    pl.table_hold.reset();  // frees the tables with the last plan sharing them
  } else {                  // plan_init failed before handing them to an owner
    for (void* p : {(void*)pl.groups, (void*)pl.mp_off, (void*)pl.tab64y, (void*)pl.tab32y, (void*)pl.otf_rc,
                    (void*)pl.tab64, (void*)pl.rot_top, (void*)pl.rot_exp, (void*)pl.rot_AB,
                    (void*)pl.rot_iB, (void*)pl.spread_tw, (void*)pl.spread_ipsi})
      if (p) (void)hipFree(p);
  }
  pl = Plan{};
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
    else prev = -1;
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

static int harvest(Plan& pl, int set) {
  if (!pl.pending[set]) return DSHT_OK;
  hipEvent_t* ev = &pl.events[(size_t)set * (DSHT_NSTAGES + 1)];
  DSHT_HIP(hipEventSynchronize(ev[DSHT_NSTAGES]));
  for (int k = 0; k < DSHT_NSTAGES; ++k) {
    float ms = 0.f;
    DSHT_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
    pl.stage_ms[k] += ms;
  }
  pl.timed += 1;
  pl.pending[set] = 0;
  return DSHT_OK;
}

// One (batched) transform.  Either `shared`: nsets weight rows
// wsets[set * ldw + i] on one point set (theta[0], phi[0], Np[0]); or nsets
// independent point sets (theta[set], phi[set], wt[set], Np[set]).  Output
// alm[set][Nlm][2], status[set] (may be NULL).
struct BatchArgs {
  int nsets = 1;
  bool shared = false;
  const int64_t* Np = nullptr;
  const double* const* theta = nullptr;
  const double* const* phi = nullptr;
  const double* const* wt = nullptr;
  const double* wsets = nullptr;
  int64_t ldw = 0;
};

static int run_batch(Plan& pl, const BatchArgs& a, double* alm, int* status, int mode,
                     hipStream_t s) {
  if (a.nsets < 1 || a.nsets > kMaxSets) {
    set_error("dsht: need 1 <= nsets <= " + std::to_string(kMaxSets));
    return DSHT_EINVAL;
  }
  if (!alm || !a.Np || !a.theta || !a.phi) {
    set_error("dsht_transform: null argument");
    return DSHT_EINVAL;
  }
  const int npts_sets = a.shared ? 1 : a.nsets;
  int64_t maxNp = 0;
  for (int k = 0; k < npts_sets; ++k) {
    const int64_t n = a.Np[k];
    if (n < 0 || n >= (int64_t)1 << 31) {
      set_error("dsht_transform: need 0 <= Np < 2^31");
      return DSHT_EINVAL;
    }
    const bool wok = a.shared ? (a.wsets != nullptr && a.ldw >= n) : (a.wt && a.wt[k]);
    if (n > 0 && (!a.theta[k] || !a.phi[k] || !wok)) {
      set_error("dsht_transform: null input pointer or ldw < Np");
      return DSHT_EINVAL;
    }
    maxNp = std::max(maxNp, n);
  }
  // the scatter engine: of the shared points for weight sets, the matrix
  // engine for grouped point sets, per set otherwise (set again below)
  pl.scatter_now = a.shared ? pick_scatter(pl, a.Np[0],
                                            a.nsets >= kSetsScatterMin ? kPickWeightSets
                                                                       : kPickPointSet,
                                            a.nsets)
                           : pick_scatter(pl, maxNp);
  const bool sets_scatter = a.shared && uses_sets_scatter(pl, a.nsets);
  const bool many = !a.shared && uses_many(pl, a.nsets, mode);
  if (many) pl.scatter_now = DSHT_SCATTER_MATRIX;
  const bool mfma = use_mfma_contract(pl, a.nsets);
  // weight sets on the spreading scatter (matrix-core contraction): on the
  // node variant each set is spread and folded straight into its slot;
  // otherwise each set is spread into a partial slot of its own and the
  // slots are folded a group at a time, so the fold writes whole rows of 8
  // sets (one set at a time writes 64 of every 16 * 4 * nsets bytes of a
  // coefficient row)
  const bool node_sets = a.shared && mfma && uses_spread(pl) && node_spread_fits(pl);
  const int spread_group = a.shared && a.nsets > 1 && mfma && uses_spread(pl) && !node_sets
                               ? std::min(a.nsets, kSpreadFoldGroup)
                               : 0;
  std::vector<ManySets> groups;
  std::vector<int> group_first, dense_sets;
  // point sets on the automatic engine: a set dense enough for the spreading
  // scatter goes alone through it (its cost is per point and per run, the
  // matrix engine's per point and order); the others are sorted, scattered
  // (matrix engine) and folded in groups of consecutive sets
  auto dense = [&](int k) {
    return pl.scatter == DSHT_SCATTER_AUTO && pick_scatter(pl, a.Np[k]) == DSHT_SCATTER_SPREAD;
  };
  int rc = DSHT_OK;
  if (many) {
    WsSizes z{1, 1, 1, 1};
    int maxg = 1;
    int64_t max_dense = 0;
    const int cap = many_group_cap(pl);
    for (int g0 = 0; g0 < a.nsets;) {
      if (dense(g0)) {
        dense_sets.push_back(g0);
        max_dense = std::max(max_dense, a.Np[g0++]);
        continue;
      }
      ManySets ms{};
      int64_t tot = a.Np[g0];
      int g = 1;
      while (g0 + g < a.nsets && g < cap && !dense(g0 + g) && tot + a.Np[g0 + g] <= kManyPoints)
        tot += a.Np[g0 + g++];
      ms.nsets = g;
      const int tile = sort_tile(tot);
      const int chunk = work_chunk(pl.chunk_max, tot);
      int64_t runs = 0;
      for (int k = 0; k < g; ++k) {
        const int n = (int)a.Np[g0 + k];
        ms.th[k] = a.theta[g0 + k];
        ms.ph[k] = a.phi[g0 + k];
        ms.w[k] = a.wt[g0 + k];
        ms.n[k] = n;
        ms.tile[k] = tile;
        ms.chunk[k] = chunk;
        ms.off[k + 1] = ms.off[k] + n;
        ms.tile_off[k + 1] = ms.tile_off[k] + (n + tile - 1) / tile;
        ms.chunk_off[k + 1] = ms.chunk_off[k] + (n + chunk - 1) / chunk;
        runs += std::min(n, pl.nseg + 2);
      }
      z.points = std::max<int64_t>(z.points, ms.off[g]);
      z.tiles = std::max<int64_t>(z.tiles, ms.tile_off[g]);
      z.chunks = std::max<int64_t>(z.chunks, ms.chunk_off[g]);
      z.runs = std::max<int64_t>(z.runs, ms.chunk_off[g] + runs);
      maxg = std::max(maxg, g);
      groups.push_back(ms);
      group_first.push_back(g0);
      g0 += g;
    }
    if (!groups.empty()) rc = ensure_workspace_z(pl, z, a.nsets, false, 1, maxg);
    if (rc == DSHT_OK && max_dense > 0) rc = ensure_workspace(pl, max_dense, a.nsets);
  } else {
    rc = ensure_workspace(pl, maxNp, a.nsets, a.shared && !scatter_gathers(pl),
                          sets_scatter ? a.nsets : std::max(1, spread_group),
                          node_sets ? a.nsets : 0);
  }
  if (rc) return rc;
  const float* full32 = nullptr;
  if (mfma && (rc = ensure_full32(pl, &full32))) return rc;
  Workspace& ws = pl.ws;
  hipEvent_t* ev = nullptr;
  if (pl.timing && a.nsets == 1 && !a.shared) {
    const int set = pl.next_set;
    pl.next_set = (pl.next_set + 1) % Plan::kTimingSets;
    if ((rc = harvest(pl, set))) return rc;
    ev = &pl.events[(size_t)set * (DSHT_NSTAGES + 1)];
    pl.pending[set] = 1;
  }
  auto mark = [&](int k) -> int {
    if (ev) DSHT_HIP(hipEventRecord(ev[k], s));
    return DSHT_OK;
  };
  if ((rc = mark(0))) return rc;
  if (a.shared) {
    const int64_t Np = a.Np[0];
    // the weight-set scatter takes longer work items: each run (a stretch of
    // one segment inside a work item) writes nsets partials that the fold
    // reads back, and there are about one run per work item plus one per
    // segment (fewer items: smaller capacities than the workspace holds)
    pl.chunk = work_chunk(sets_scatter ? std::max(pl.chunk_max, kSetsChunk) : chunk_cap(pl), Np);
    ws.cur_chunks = (int)((Np + pl.chunk - 1) / pl.chunk);
    if ((rc = launch_points(pl, Np, a.theta[0], a.phi[0], nullptr, mode, ws.flags, s))) return rc;
    const bool gathers = scatter_gathers(pl);
    if (sets_scatter && (rc = launch_scatter_sets(pl, Np, a.wsets, a.ldw, a.nsets, s))) return rc;
    // one slot per set: fold every set in one launch
    const bool fold_all = sets_scatter && mfma;
    if (fold_all && (rc = launch_fold_sets(pl, a.nsets, s, ws.part, part_set_stride(pl)))) return rc;
    if (node_sets) {  // point-major FP32 weight rows, then every set in one launch
      if ((rc = launch_place_weight_sets(pl, Np, a.wsets, a.ldw, a.nsets, s))) return rc;
      if ((rc = launch_spread_nodes_sets(pl, 0, a.nsets, a.nsets, true, s))) return rc;
    }
    for (int g0 = 0; spread_group && g0 < a.nsets; g0 += spread_group) {
      const int g = std::min(spread_group, a.nsets - g0);
      for (int k = 0; k < g; ++k) {
        const double* w = a.wsets + (g0 + k) * a.ldw;
        float* part = (float*)ws.part + k * part_set_stride(pl);
        if ((rc = launch_weight_max(pl, Np, w, s, ws.wmax + g0 + k))) return rc;
        if ((rc = launch_scatter_spread(pl, Np, w, 2, s, part, ws.wmax + g0 + k))) return rc;
      }
      if ((rc = launch_fold_sets(pl, a.nsets, s, ws.part, part_set_stride(pl), g0, g))) return rc;
    }
    for (int set = 0; set < a.nsets && !fold_all && !spread_group && !node_sets; ++set) {
      const double* w = a.wsets + set * a.ldw;
      if (sets_scatter) {
        const float* part = (const float*)ws.part + set * part_set_stride(pl);
        if ((rc = launch_fold(pl, a.nsets, set, s, part))) return rc;
        if (!mfma && (rc = launch_contract_set(pl, a.nsets, set, s))) return rc;
        continue;
      }
      if (gathers && uses_node_spread(pl, a.nsets)) {  // scatter and fold in one launch
        if ((rc = launch_spread_nodes(pl, Np, w, 2, s))) return rc;
        if (!mfma && (rc = launch_contract_set(pl, a.nsets, set, s))) return rc;
        continue;
      }
      if (gathers) {
        rc = uses_spread(pl) ? launch_scatter_spread(pl, Np, w, 2, s)
                             : launch_scatter_mfma(pl, Np, w, true, s);
        if (rc) return rc;
      } else {
        if ((rc = launch_gather_weights(pl, Np, w, s))) return rc;
        if ((rc = launch_scatter(pl, Np, ws.wsort, s))) return rc;
      }
      if ((rc = launch_fold(pl, a.nsets, set, s))) return rc;
      if (!mfma && (rc = launch_contract_set(pl, a.nsets, set, s))) return rc;
    }
  } else if (many) {
    pl.scatter_now = DSHT_SCATTER_MATRIX;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      const ManySets& ms = groups[gi];
      const int g0 = group_first[gi];
      if ((rc = launch_points_many(pl, ms, mode, ws.flags + 4 * g0, ws.wmax + g0, s))) return rc;
      if ((rc = launch_scatter_mfma_many(pl, ms, ws.wmax + g0, s))) return rc;
      if ((rc = launch_fold_many(pl, ms, g0, a.nsets, s))) return rc;
    }
    // the dense sets after every group: each uses the single transform's
    // arrays (ws.wmax[0] among them), which the groups' launches have read
    for (const int set : dense_sets) {
      const int64_t Np = a.Np[set];
      pl.scatter_now = DSHT_SCATTER_SPREAD;
      pl.chunk = work_chunk(chunk_cap(pl), Np);
      ws.cur_chunks = (int)((Np + pl.chunk - 1) / pl.chunk);
      if ((rc = launch_points(pl, Np, a.theta[set], a.phi[set], a.wt[set], mode, ws.flags + 4 * set, s)))
        return rc;
      if (node_spread_fits(pl)) {  // scatter and fold into the set's slot in one launch
        if ((rc = launch_spread_nodes_sets(pl, set, 1, a.nsets, false, s))) return rc;
        continue;
      }
      if ((rc = launch_scatter(pl, Np, nullptr, s))) return rc;
      if ((rc = launch_fold(pl, a.nsets, set, s))) return rc;
    }
  } else {
    for (int set = 0; set < a.nsets; ++set) {
      const int64_t Np = a.Np[set];
      pl.scatter_now = pick_scatter(pl, Np, mfma ? kPickPointSet : kPickPlain);
      pl.chunk = work_chunk(chunk_cap(pl), Np);
      ws.cur_chunks = (int)((Np + pl.chunk - 1) / pl.chunk);
      if ((rc = launch_points(pl, Np, a.theta[set], a.phi[set], a.wt[set], mode, ws.flags + 4 * set, s)))
        return rc;
      if ((rc = mark(1))) return rc;
      if (uses_node_spread(pl, a.nsets)) {  // scatter and fold in one launch
        if ((rc = launch_spread_nodes(pl, Np, nullptr, 0, s))) return rc;
        if ((rc = mark(2))) return rc;
      } else {
        if ((rc = launch_scatter(pl, Np, nullptr, s))) return rc;
        if ((rc = mark(2))) return rc;
        if ((rc = launch_fold(pl, a.nsets, set, s))) return rc;
      }
      if ((rc = mark(3))) return rc;
      if (!mfma && (rc = launch_contract_set(pl, a.nsets, set, s))) return rc;
    }
  }
  // the plain contraction ran per set, right after its fold (one coefficient
  // slot); the matrix-core contraction takes all slots at once
  if (mfma && (rc = launch_contract_mfma(pl, full32, a.nsets, s))) return rc;
  if ((rc = mark(4))) return rc;
  if ((rc = launch_rotate_combine(pl, a.nsets, ws.flags, a.shared ? 0 : 4, alm, status, s))) return rc;
  if ((rc = mark(5))) return rc;
  return DSHT_OK;
}

static int run_transform(Plan& pl, int64_t Np, const double* theta, const double* phi,
                         const double* wt, double* alm, int* status, int mode, hipStream_t s) {
  BatchArgs a;
  a.nsets = 1;
  a.Np = &Np;
  a.theta = &theta;
  a.phi = &phi;
  a.wt = &wt;
  if (Np == 0) {  // no points: inputs may be NULL
    static const double dummy = 0.0;
    static const double* dp = &dummy;
    a.theta = a.phi = a.wt = &dp;
  }
  return run_batch(pl, a, alm, status, mode, s);
}

}  // namespace dsht

struct dsht_plan {
  dsht::Plan pl;
};

using namespace dsht;

extern "C" {

const char* dsht_last_error(void) { return g_err.c_str(); }

int dsht_plan_engine(const dsht_plan* p, int64_t Np, int* engine) {
  if (!p || !engine || Np < 0) {
    set_error("dsht_plan_engine: null argument or Np < 0");
    return DSHT_EINVAL;
  }
  const Plan& pl = p->pl;
  if (pl.precision != DSHT_FP32) {
    *engine = DSHT_SCATTER_VECTOR;
    return DSHT_OK;
  }
  int e = pick_scatter(pl, Np, kPickPlain);
  if (e == DSHT_SCATTER_SPREAD && pl.Nl > kSpreadMaxNl) e = DSHT_SCATTER_MATRIX;
  *engine = e;
  return DSHT_OK;
}

int dsht_plan_configure(dsht_plan* p, int option, int value) {
  if (!p) {
    set_error("dsht_plan_configure: null plan");
    return DSHT_EINVAL;
  }
  Plan& pl = p->pl;
  if (option == DSHT_OPT_SCATTER) {
    if (value != DSHT_SCATTER_MATRIX && value != DSHT_SCATTER_VECTOR &&
        value != DSHT_SCATTER_SPREAD && value != DSHT_SCATTER_AUTO) {
      set_error("dsht_plan_configure: unknown scatter engine");
      return DSHT_EINVAL;
    }
    pl.scatter = value;
    return DSHT_OK;
  }
  if (option == DSHT_OPT_CONTRACT) {
    if (value != DSHT_CONTRACT_MATRIX && value != DSHT_CONTRACT_VECTOR) {
      set_error("dsht_plan_configure: unknown contraction engine");
      return DSHT_EINVAL;
    }
    pl.contract = value;
    return DSHT_OK;
  }
  if (option == DSHT_OPT_CHUNK) {
    if (value < 64 || value > 65536) {
      set_error("dsht_plan_configure: chunk must be in [64, 65536]");
      return DSHT_EINVAL;
    }
    if (value != pl.chunk_max) {
      DeviceGuard g(pl.device);
      DSHT_HIP(hipDeviceSynchronize());  // queued transforms use the old decomposition
      pl.chunk_max = pl.chunk = value;
      pl.ws.cap_chunks = pl.ws.cap_runs = 0;  // force a workspace resize on next use
    }
    pl.chunk_explicit = true;
    return DSHT_OK;
  }
  set_error("dsht_plan_configure: unknown option");
  return DSHT_EINVAL;
}

int dsht_alm2cl(int Nl, int nsets, const double* alm, int64_t ld, double* cl, int cross,
                void* stream) {
  const int64_t Nlm = (int64_t)Nl * (Nl + 1) / 2;
  if (Nl < 1 || nsets < 1 || nsets > kMaxSets || ld < Nlm || !alm || !cl) {
    set_error("dsht_alm2cl: need Nl >= 1, 1 <= nsets <= 64, ld >= Nl(Nl+1)/2, non-null arrays");
    return DSHT_EINVAL;
  }
  return launch_alm2cl(Nl, nsets, alm, ld, cl, cross, (hipStream_t)stream);
}

int dsht_threej000_table(int N, double* store, void* stream) {
  if (N < 1 || N > tj_max_n() || !store) {
    set_error("dsht_threej000_table: need 1 <= N <= " + std::to_string(tj_max_n()) +
              " and a non-null store");
    return DSHT_EINVAL;
  }
  return launch_threej000_table(N, store, (hipStream_t)stream);
}

int dsht_mode_coupling(int Nl, const double* W, int64_t nW, double* M, void* stream) {
  if (Nl < 1 || Nl > tj_max_n() || nW < 0 || (nW > 0 && !W) || !M) {
    set_error("dsht_mode_coupling: need 1 <= Nl <= " + std::to_string(tj_max_n()) +
              ", nW >= 0, W non-null if nW > 0, and a non-null M");
    return DSHT_EINVAL;
  }
  return launch_mode_coupling(Nl, W, nW, M, (hipStream_t)stream);
}

int dsht_synthesis_rings(int Nl, const double* alm, int nrings, const double* x, double* F,
                         void* stream) {
  if (Nl < 1 || Nl > synth_max_nl() || nrings < 0 || nrings > synth_max_rings() ||
      (nrings > 0 && (!alm || !x || !F))) {
    set_error("dsht_synthesis_rings: need 1 <= Nl <= " + std::to_string(synth_max_nl()) +
              ", 0 <= nrings <= " + std::to_string(synth_max_rings()) + " and non-null arrays");
    return DSHT_EINVAL;
  }
  if (nrings == 0) return DSHT_OK;
  return launch_synth_rings(Nl, alm, nrings, x, F, (hipStream_t)stream);
}

int dsht_plan_debug_rhs(dsht_plan* p, void* out, int64_t bytes) {
  if (!p || !out) {
    set_error("dsht_plan_debug_rhs: null argument");
    return DSHT_EINVAL;
  }
  Plan& pl = p->pl;
  const int64_t need = (int64_t)pl.Nl * 2 * pl.Npair * 24 * (pl.precision == DSHT_FP32 ? 4 : 8);
  if (bytes < need || !pl.ws.rhs) {
    set_error("dsht_plan_debug_rhs: buffer too small or no transform yet");
    return DSHT_EINVAL;
  }
  DeviceGuard g(pl.device);
  DSHT_HIP(hipDeviceSynchronize());
  DSHT_HIP(hipMemcpy(out, pl.ws.rhs, need, hipMemcpyDeviceToHost));
  return DSHT_OK;
}

int dsht_plan_set_timing(dsht_plan* p, int enable) {
  if (!p) {
    set_error("dsht_plan_set_timing: null plan");
    return DSHT_EINVAL;
  }
  Plan& pl = p->pl;
  DeviceGuard g(pl.device);
  if (enable && pl.events.empty()) {
    pl.events.resize((size_t)Plan::kTimingSets * (DSHT_NSTAGES + 1));
    for (hipEvent_t& e : pl.events) DSHT_HIP(hipEventCreate(&e));
    pl.pending.assign(Plan::kTimingSets, 0);
  }
  for (int k = 0; k < (int)pl.pending.size(); ++k) {
    if (pl.pending[k]) DSHT_HIP(hipEventSynchronize(pl.events[(size_t)k * (DSHT_NSTAGES + 1) + DSHT_NSTAGES]));
    pl.pending[k] = 0;
  }
  for (double& v : pl.stage_ms) v = 0.0;
  pl.timed = 0;
  pl.timing = enable != 0;
  return DSHT_OK;
}

int dsht_plan_get_timing(dsht_plan* p, double stage_ms[DSHT_NSTAGES], int64_t* ntransforms) {
  if (!p || !stage_ms) {
    set_error("dsht_plan_get_timing: null argument");
    return DSHT_EINVAL;
  }
  Plan& pl = p->pl;
  DeviceGuard g(pl.device);
  for (int k = 0; k < (int)pl.pending.size(); ++k) {
    int rc = harvest(pl, k);
    if (rc) return rc;
  }
  for (int k = 0; k < DSHT_NSTAGES; ++k) stage_ms[k] = pl.stage_ms[k];
  if (ntransforms) *ntransforms = pl.timed;
  return DSHT_OK;
}

int dsht_plan_create(int Nl, int Nx, double xmax, int precision, dsht_plan** out) {
  if (!out) {
    set_error("dsht_plan_create: null output");
    return DSHT_EINVAL;
  }
  *out = nullptr;
  dsht_plan* p = new dsht_plan();
  int rc = plan_init(p->pl, Nl, Nx, xmax, precision, true, true);
  if (rc) {
    plan_free(p->pl);
    delete p;
    return rc;
  }
  *out = p;
  return DSHT_OK;
}

int dsht_plan_create_from_tables(int Nl, int Nx, double xmax, int precision, const double* Yv,
                                 dsht_plan** out) {
  if (!out || !Yv) {
    set_error("dsht_plan_create_from_tables: null argument");
    return DSHT_EINVAL;
  }
  *out = nullptr;
  dsht_plan* p = new dsht_plan();
  Plan& pl = p->pl;
  int rc = plan_init(pl, Nl, Nx, xmax, precision, true, false);
  // the rows go through a device slab of at most 256 MiB (at least one row)
  double* slab = nullptr;
  const int64_t per = std::max<int64_t>(1, ((int64_t)256 << 20) / ((int64_t)Nx * 8));
  if (rc == DSHT_OK) {
    const size_t bytes = (size_t)std::min<int64_t>(per, pl.Nlm) * Nx * sizeof(double);
    if (hipMalloc((void**)&slab, bytes) != hipSuccess) {
      slab = nullptr;
      set_error("dsht_plan_create_from_tables: out of device memory for the staging slab");
      rc = DSHT_ENOMEM;
    }
  }
  // an FP32 plan holds no Y table (its contraction regenerates Y): nothing to load
  for (int64_t r0 = 0; rc == DSHT_OK && !pl.otf && r0 < pl.Nlm; r0 += per) {
    const int64_t n = std::min<int64_t>(per, pl.Nlm - r0);
    hipError_t e = hipMemcpy(slab, Yv + r0 * Nx, (size_t)n * Nx * sizeof(double),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_error(std::string("dsht_plan_create_from_tables: ") + hipGetErrorString(e));
      rc = DSHT_EHIP;
      break;
    }
    rc = launch_relayout_y(pl, r0, n, slab, 0);
    if (rc == DSHT_OK && hipDeviceSynchronize() != hipSuccess) {  // the slab is reused
      set_error("dsht_plan_create_from_tables: relayout failed");
      rc = DSHT_EHIP;
    }
  }
  if (slab) (void)hipFree(slab);
  if (rc) {
    plan_free(pl);
    delete p;
    return rc;
  }
  *out = p;
  return DSHT_OK;
}

int dsht_plan_destroy(dsht_plan* p) {
  if (!p) return DSHT_OK;
  DeviceGuard g(p->pl.device);
  (void)hipDeviceSynchronize();
  plan_free(p->pl);
  delete p;
  return DSHT_OK;
}

int dsht_plan_clone(const dsht_plan* src, dsht_plan** out) {
  if (!src || !out) {
    set_error("dsht_plan_clone: null argument");
    return DSHT_EINVAL;
  }
  dsht_plan* p = new dsht_plan();
  Plan& d = p->pl;
  const Plan& s = src->pl;
  d.device = s.device;
  d.Nl = s.Nl;
  d.Nx = s.Nx;
  d.Npair = s.Npair;
  d.nseg = s.nseg;
  d.Nlm = s.Nlm;
  d.xmax = s.xmax;
  d.dx = s.dx;
  d.acos_xmax = s.acos_xmax;
  d.acos_mxmax = s.acos_mxmax;
  d.precision = s.precision;
  d.chunk = s.chunk;
  d.chunk_max = s.chunk_max;
  d.chunk_explicit = s.chunk_explicit;
  d.scatter = s.scatter;
  d.contract = s.contract;
  d.ngroups = s.ngroups;
  d.groups = s.groups;
  d.nmgroups = s.nmgroups;
  d.mgroups = s.mgroups;
  d.mp_off = s.mp_off;
  d.tab_vec4 = s.tab_vec4;
  d.tab64y = s.tab64y;
  d.tab32y = s.tab32y;
  d.otf = s.otf;
  d.otf_rc = s.otf_rc;
  d.tab64 = s.tab64;
  d.caller_tables = s.caller_tables;
  d.table_bytes = s.table_bytes;
  d.rot_top = s.rot_top;
  d.rot_exp = s.rot_exp;
  d.rot_AB = s.rot_AB;
  d.rot_iB = s.rot_iB;
  d.spread_logn = s.spread_logn;
  d.spread_beta = s.spread_beta;
  d.spread_tw = s.spread_tw;
  d.spread_ipsi = s.spread_ipsi;
  d.table_hold = s.table_hold;
  *out = p;
  return DSHT_OK;
}

int dsht_plan_reserve_for(dsht_plan* p, int64_t Np, int nsets, int kinds) {
  if (!p || Np < 0 || Np >= (int64_t)1 << 31 || nsets < 1 || nsets > kMaxSets ||
      kinds < 1 || kinds > (DSHT_RESERVE_POINT_SETS | DSHT_RESERVE_WEIGHT_SETS)) {
    set_error("dsht_plan_reserve: null plan, Np outside [0, 2^31), nsets outside [1, 64] "
              "or kinds not a non-empty combination of DSHT_RESERVE_*");
    return DSHT_EINVAL;
  }
  DeviceGuard g(p->pl.device);
  Plan& pl = p->pl;
  int rc = DSHT_OK;
  if (kinds & DSHT_RESERVE_WEIGHT_SETS) {
    // room for either engine a transform of up to Np points may take
    // (the node spreading of weight sets: their point-major rows)
    pl.scatter_now = pick_scatter(pl, Np, kPickWeightSets, nsets);
    const bool rows = use_mfma_contract(pl, nsets) && uses_spread(pl) && node_spread_fits(pl);
    pl.scatter_now = pick_scatter(pl, 0);
    const bool sets = uses_sets_scatter(pl, nsets);
    const bool wsort = nsets > 1 && !scatter_gathers(pl);
    rc = ensure_workspace(pl, Np, nsets, wsort, sets ? nsets : 1, rows ? nsets : 0);
  }
  if (rc == DSHT_OK && (kinds & DSHT_RESERVE_POINT_SETS)) {
    rc = ensure_workspace(pl, Np, nsets);
    // point-set batches (dsht_transform_many) sort groups of sets together
    if (rc == DSHT_OK && uses_many(pl, nsets, DSHT_MODE_CALL))
      rc = ensure_workspace_z(pl, many_sizes(pl, Np, nsets), nsets, false, 1,
                              std::min(nsets, many_group_cap(pl)));
  }
  if (rc == DSHT_OK && use_mfma_contract(p->pl, nsets)) {
    const float* full32 = nullptr;
    rc = ensure_full32(p->pl, &full32);  // so that later transforms allocate nothing
  }
  return rc;
}

int dsht_plan_reserve(dsht_plan* p, int64_t Np, int nsets) {
  return dsht_plan_reserve_for(p, Np, nsets, DSHT_RESERVE_POINT_SETS | DSHT_RESERVE_WEIGHT_SETS);
}

int dsht_plan_footprint(const dsht_plan* p, int64_t* bytes) {
  if (!p || !bytes) {
    set_error("dsht_plan_footprint: null argument");
    return DSHT_EINVAL;
  }
  int64_t full = 0;
  if (p->pl.table_hold) {
    std::lock_guard<std::mutex> lock(p->pl.table_hold->mu);
    full = p->pl.table_hold->full32_bytes;
  }
  *bytes = p->pl.table_bytes + full + p->pl.ws.bytes;
  return DSHT_OK;
}

int dsht_plan_estimate(int Nl, int Nx, int precision, int64_t Np, int64_t* table_bytes,
                       int64_t* work_bytes) {
  if (Nl < 2 || Nx < 2 || Nl > kMaxNl || Nx > kMaxNx || Np < 0) {
    set_error("dsht_plan_estimate: need 2 <= Nl <= " + std::to_string(kMaxNl) + ", 2 <= Nx <= " +
              std::to_string(kMaxNx) + " and Np >= 0");
    return DSHT_EINVAL;
  }
  std::vector<GroupDesc> groups;
  std::vector<int64_t> mp_off;
  int64_t vec4 = 0;
  const int Npair = node_pairs(Nx);
  build_layout(Nl, Npair, groups, mp_off, vec4);
  const size_t nm = xcd_schedule(pair_groups(groups), kXcds).size();
  groups = xcd_schedule(groups, kXcds);
  const int64_t Nlm = (int64_t)Nl * (Nl + 1) / 2;
  // FP32 plans keep only the recurrence's coefficients (no Y table)
  const int64_t tab = precision == DSHT_FP32 ? 16 * otf_offset(Nl, Nl) : vec4 * 2 * 8;
  if (table_bytes)
    *table_bytes = tab + Nlm * (3 * 8 + 4) + (int64_t)(groups.size() + nm) * sizeof(GroupDesc) +
                   (int64_t)mp_off.size() * 8;
  if (work_bytes) *work_bytes = workspace_bytes(Nl, Nx, precision, Np, 2048);
  return DSHT_OK;
}

int dsht_plan_get_tables(dsht_plan* p, double* Yv, double* Yd) {
  if (!p) {
    set_error("dsht_plan_get_tables: null plan");
    return DSHT_EINVAL;
  }
  Plan& pl = p->pl;
  DeviceGuard g(pl.device);
  const size_t bytes = (size_t)pl.Nlm * pl.Nx * sizeof(double);
  double *dv = nullptr, *dd = nullptr;
  if (hipMalloc((void**)&dv, bytes) != hipSuccess || hipMalloc((void**)&dd, bytes) != hipSuccess) {
    if (dv) (void)hipFree(dv);
    set_error("dsht_plan_get_tables: out of device memory");
    return DSHT_ENOMEM;
  }
  int rc = launch_tables(pl, 3, dv, dd, 0);
  if (rc == DSHT_OK) {
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && Yv) e = hipMemcpy(Yv, dv, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess && Yd) e = hipMemcpy(Yd, dd, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_error(std::string("dsht_plan_get_tables: ") + hipGetErrorString(e));
      rc = DSHT_EHIP;
    }
  }
  (void)hipFree(dv);
  (void)hipFree(dd);
  return rc;
}

int dsht_transform(dsht_plan* p, int64_t Np, const double* theta, const double* phi,
                   const double* wt, double* alm, int* status, int mode, void* stream) {
  if (!p) {
    set_error("dsht_transform: null plan");
    return DSHT_EINVAL;
  }
  if (mode != DSHT_MODE_CALL && mode != DSHT_MODE_BASIC) {
    set_error("dsht_transform: bad mode");
    return DSHT_EINVAL;
  }
  DeviceGuard g(p->pl.device);
  return run_transform(p->pl, Np, theta, phi, wt, alm, status, mode, (hipStream_t)stream);
}

int dsht_transform_noalloc(dsht_plan* p, int64_t Np, const double* theta, const double* phi,
                           const double* wt, double* alm, int* status, int mode, void* stream) {
  if (!p) {
    set_error("dsht_transform_noalloc: null plan");
    return DSHT_EINVAL;
  }
  if (mode != DSHT_MODE_CALL && mode != DSHT_MODE_BASIC) {
    set_error("dsht_transform_noalloc: bad mode");
    return DSHT_EINVAL;
  }
  if (Np < 0 || Np >= (int64_t)1 << 31) {
    set_error("dsht_transform_noalloc: need 0 <= Np < 2^31");
    return DSHT_EINVAL;
  }
  Plan& pl = p->pl;
  DeviceGuard g(pl.device);
  char* base = nullptr;
  int rc = scratch_address(pl.device, &base);
  if (rc) return rc;
  Workspace fixed;
  int64_t need = 0;
  if (carve_workspace(pl, Np, base, kScratchBytes, fixed, &need) != DSHT_OK) {
    set_error("dsht_transform_noalloc: a transform of " + std::to_string(Np) + " points at Nl=" +
              std::to_string(pl.Nl) + ", Nx=" + std::to_string(pl.Nx) + " needs " +
              std::to_string(need) + " bytes of scratch; the fixed buffer holds " +
              std::to_string(kScratchBytes) + " (use dsht_transform)");
    return DSHT_EINVAL;
  }
  // the plan's own workspace is set aside (not freed, not grown) for the call
  const Workspace own = pl.ws;
  pl.ws = fixed;
  rc = run_transform(pl, Np, theta, phi, wt, alm, status, mode, (hipStream_t)stream);
  pl.ws = own;
  return rc;
}

int dsht_transform_sets(dsht_plan* p, int64_t Np, const double* theta, const double* phi,
                        int nsets, const double* wt, int64_t ldw, double* alm, int* status,
                        int mode, void* stream) {
  if (!p) {
    set_error("dsht_transform_sets: null plan");
    return DSHT_EINVAL;
  }
  if (mode != DSHT_MODE_CALL && mode != DSHT_MODE_BASIC) {
    set_error("dsht_transform_sets: bad mode");
    return DSHT_EINVAL;
  }
  BatchArgs a;
  a.nsets = nsets;
  a.shared = true;
  a.Np = &Np;
  a.theta = &theta;
  a.phi = &phi;
  a.wsets = wt;
  a.ldw = ldw;
  DeviceGuard g(p->pl.device);
  return run_batch(p->pl, a, alm, status, mode, (hipStream_t)stream);
}

int dsht_transform_many(dsht_plan* p, int nsets, const int64_t* Np, const double* const* theta,
                        const double* const* phi, const double* const* wt, double* alm,
                        int* status, int mode, void* stream) {
  if (!p) {
    set_error("dsht_transform_many: null plan");
    return DSHT_EINVAL;
  }
  if (mode != DSHT_MODE_CALL && mode != DSHT_MODE_BASIC) {
    set_error("dsht_transform_many: bad mode");
    return DSHT_EINVAL;
  }
  if (!wt) {
    set_error("dsht_transform_many: null weight array");
    return DSHT_EINVAL;
  }
  BatchArgs a;
  a.nsets = nsets;
  a.Np = Np;
  a.theta = theta;
  a.phi = phi;
  a.wt = wt;
  DeviceGuard g(p->pl.device);
  return run_batch(p->pl, a, alm, status, mode, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Reference (csht) C ABI
// ---------------------------------------------------------------------------

int make_table(int Nl, int Nx, double xmax, double Yv[], double Yd[]) {
  Plan pl;
  int rc = plan_init(pl, Nl, Nx, xmax, DSHT_FP64, false, false);
  if (rc == DSHT_OK) {
    const size_t bytes = (size_t)pl.Nlm * Nx * sizeof(double);
    double *dv = nullptr, *dd = nullptr;
    if (hipMalloc((void**)&dv, bytes) != hipSuccess || hipMalloc((void**)&dd, bytes) != hipSuccess) {
      set_error("make_table: out of device memory");
      rc = DSHT_ENOMEM;
    } else {
      rc = launch_tables(pl, 3, dv, dd, 0);
      if (rc == DSHT_OK) {
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(Yv, dv, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(Yd, dd, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
          set_error("make_table: device copy failed");
          rc = DSHT_EHIP;
        }
      }
    }
    if (dv) (void)hipFree(dv);
    if (dd) (void)hipFree(dd);
  }
  plan_free(pl);
  return rc == DSHT_OK ? 0 : 1;
}

int do_transform(int Nl, int Nx, double xmax, double Yv[], double Yd[], int Np, double cost[],
                 double phi[], double wt[], double carr[], double sarr[]) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  Plan pl;
  int rc = plan_init(pl, Nl, Nx, xmax, DSHT_FP64, true, false, true);
  const size_t tb = (size_t)pl.Nlm * Nx * sizeof(double);
  double *dv = nullptr, *dd = nullptr, *dc = nullptr, *dp = nullptr, *dw = nullptr, *da = nullptr;
  std::vector<double> alm;
  if (rc == DSHT_OK) {
    const size_t pb = std::max<size_t>(1, (size_t)Np) * sizeof(double);
    if (hipMalloc((void**)&dv, tb) != hipSuccess || hipMalloc((void**)&dd, tb) != hipSuccess ||
        hipMalloc((void**)&dc, pb) != hipSuccess || hipMalloc((void**)&dp, pb) != hipSuccess ||
        hipMalloc((void**)&dw, pb) != hipSuccess ||
        hipMalloc((void**)&da, pl.Nlm * 2 * sizeof(double)) != hipSuccess) {
      set_error("do_transform: out of device memory");
      rc = DSHT_ENOMEM;
    }
  }
  if (rc == DSHT_OK) {
    // the caller's tables, not a rebuild: same semantics as csht
    if (hipMemcpy(dv, Yv, tb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dd, Yd, tb, hipMemcpyHostToDevice) != hipSuccess ||
        (Np > 0 && (hipMemcpy(dc, cost, Np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(dp, phi, Np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(dw, wt, Np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))) {
      set_error("do_transform: host to device copy failed");
      rc = DSHT_EHIP;
    }
  }
  if (rc == DSHT_OK) rc = launch_relayout(pl, dv, dd, 0);
  if (rc == DSHT_OK) rc = run_transform(pl, Np, dc, dp, dw, da, nullptr, DSHT_MODE_CSHT, 0);
  if (rc == DSHT_OK) {
    alm.resize(pl.Nlm * 2);
    if (hipMemcpy(alm.data(), da, alm.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("do_transform: device to host copy failed");
      rc = DSHT_EHIP;
    }
  }
  if (rc == DSHT_OK) {
    for (int64_t i = 0; i < pl.Nlm; ++i) {
      carr[i] = alm[2 * i];
      sarr[i] = -alm[2 * i + 1];
    }
  }
  for (double* q : {dv, dd, dc, dp, dw, da})
    if (q) (void)hipFree(q);
  plan_free(pl);
  return rc == DSHT_OK ? 0 : 1;
}

}  // extern "C"
