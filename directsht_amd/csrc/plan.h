// Plan: resident tables + per-transform workspace of one DirectSHT(Nl, Nx, xmax).
#pragma once

#include <memory>
#include <mutex>
#include <vector>

#include "common.h"

namespace dsht {

// Node pairs of the contraction layout: an even number covering Nx nodes
// (nodes >= Nx are padding with zero table values and coefficients), so that
// the Y-only table can hold two pairs of a row in one 16-byte vector.
inline int node_pairs(int Nx) { return 2 * ((Nx + 3) / 4); }

// Band and grid limits of transform plans: the polar rotation stages 4 Nl
// doubles per degree in LDS (128 KB at Nl = 4096); Nx bounds the segment and
// node-pair counts well inside int range.
constexpr int kMaxNl = 4096;
constexpr int kMaxNx = 1 << 20;

// Point sets of one group of a dsht_transform_many call, processed together
// (kManyGroup at most): passed by value to the group's kernels.  Set s has
// n[s] points (device arrays th, ph, w); its sorted records occupy global
// positions off[s] .. off[s] + n[s]; its sort tiles are global tiles
// tile_off[s] .. tile_off[s+1] - 1 of tile[s] points, its scatter work items
// global chunks chunk_off[s] .. chunk_off[s+1] - 1 of chunk[s] points.
constexpr int kManyGroup = 16;
struct ManySets {
  int nsets;
  const double* th[kManyGroup];
  const double* ph[kManyGroup];
  const double* w[kManyGroup];
  int n[kManyGroup];
  int tile[kManyGroup];
  int chunk[kManyGroup];
  int off[kManyGroup + 1];
  int tile_off[kManyGroup + 1];
  int chunk_off[kManyGroup + 1];
};

// the set of global index x given the sets' first indices first[0..nsets]
__host__ __device__ inline int many_set_of(const int* first, int nsets, int x) {
  int lo = 0, hi = nsets - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct Workspace {
  int64_t cap_points = 0;  // capacity of the sorted-point arrays
  int64_t cap_chunks = 0;  // capacity of per-chunk arrays
  int64_t cap_runs = 0;    // capacity of the run partials
  int64_t cap_tiles = 0;   // capacity of the tile histogram matrix (rows)
  int cap_sets = 0;        // sets of a batched transform the per-set arrays hold
  int cap_segsets = 1;     // point sets whose segment arrays are held at once (many path)
  bool has_wsort = false;  // wsort allocated (weight-set transforms)
  int cur_chunks = 0;      // scatter chunks of the transform being enqueued
  PointRec* rec = nullptr;    // [cap_points] sorted point records
  int* run_count = nullptr;   // [cap_chunks + 1]: runs per chunk
  int* run_base = nullptr;    // [cap_chunks + 1]: exclusive scan of run_count
  int* run_seg = nullptr;     // [cap_runs]: segment of each run
  void* part = nullptr;       // per-run v_k(m), k = 0..3, complex, [4][Nl][2] per
                              // run; double (FP64 scatter) or float (FP32 scatter):
                              // [cap_runs] runs, or for the weight-set scatter
                              // [cap_part_sets][cap_set_runs] (one slot per set),
                              // whichever is larger
  int cap_part_sets = 0;      // slots of part
  int cap_wsf_sets = 0;       // rows of wsf
  int64_t cap_set_points = 0; // points of a weight-set transform (wsf rows)
  int64_t cap_set_runs = 0;   // runs of a part slot of a weight-set transform
  int* seg_count = nullptr;   // [nseg + 2]
  int* seg_start = nullptr;   // [nseg + 2]
  int* seg_cursor = nullptr;  // [nseg + 2]
  int* seg_run_lo = nullptr;  // [nseg + 2]: first run of each segment (0 if empty)
  int* seg_run_hi = nullptr;  // [nseg + 2]: one past its last run (0 if empty)
  int* ne_rank = nullptr;     // [nseg + 1]: non-empty segments before s
  int* thist = nullptr;       // [cap_tiles][nseg + 1]: tile histograms, then offsets
  int2* segrank = nullptr;    // [2][cap_points] ints: each input point's segment, then its
                              // rank among its tile's points of that segment (input order)
  void* rhs = nullptr;        // node coefficients (float or double): plain contraction
                              // plain_rec (one set at a time); matrix-core
                              // contraction and caller tables [Nl][2][Npair][cap_sets][16]
  int64_t rhs_elems = 0;      // capacity of rhs in elements
  unsigned* rmax = nullptr;   // [cap_sets]: max(|c|, |d/dx|) per set (float bits)
  unsigned* wmax = nullptr;   // [cap_sets]: max |w| per weight set (float bits)
  int* sidx = nullptr;        // [cap_points]: input index of each sorted point
                              // (weight-set scatter)
  float* wsf = nullptr;       // [cap_wsf_sets][ldws]: weight sets in sorted order,
                              // FP32, zero past the valid points (weight-set scatter;
                              // the node spreading of weight sets: [point][nsets]);
                              // followed by the same size of scratch for the sort
  int64_t ldws = 0;           // row stride of wsf: cap_set_points rounded up to 64, + 64
  double* wsort = nullptr;    // [cap_points]: one weight set in sorted order
  double* a_equ = nullptr;    // [cap_sets][Nlm][2]
  double* a_pol = nullptr;    // [cap_sets][Nlm][2]
  double* a_tmp = nullptr;    // (cap_sets > 1) the batched rotation's copy of
                              // a_pol, grouped: [cap_sets / G][Nlm][2][G] in l-major
                              // order, G <= 16 (room for cap_sets rounded up to 16)
  double* a_der = nullptr;    // [kContractSplits][Nlm][8]: partial sums of the plain
                              // contraction per node range (one set at a time)
  int* flags = nullptr;       // [cap_sets][4]: status bits, n_polar, unused, unused
  int64_t bytes = 0;
  bool fixed = false;         // carved from a caller-independent fixed buffer: never
                              // grown or freed (dsht_transform_noalloc)
};

// Frees a plan's table arrays (on the plan's device) when destroyed.  Also
// holds the FP32 {Y, Y'} table of the matrix-core contraction, built on first
// use and shared by every clone of the plan.
struct TableOwner {
  int device;
  std::vector<void*> ptrs;
  std::mutex mu;             // guards full32
  float* full32 = nullptr;   // FP32 {Y, Y'} table (tab_vec4 4-vectors), or NULL
  int64_t full32_bytes = 0;
  TableOwner(int dev, std::vector<void*> p) : device(dev), ptrs(std::move(p)) {}
  TableOwner(const TableOwner&) = delete;
  TableOwner& operator=(const TableOwner&) = delete;
  ~TableOwner() {
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != device) (void)hipSetDevice(device);
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    if (full32) (void)hipFree(full32);
    if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
  }
};

struct Plan {
  int device = 0;
  int Nl = 0, Nx = 0, Npair = 0, nseg = 0;
  int64_t Nlm = 0;
  double xmax = 0, dx = 0, acos_xmax = 0, acos_mxmax = 0;
  int precision = DSHT_FP64;
  // scatter decomposition
  int chunk_max = 2048;  // configured points per scatter work item (the largest used)
  bool chunk_explicit = false;  // chunk_max set by dsht_plan_configure
  int chunk = 2048;      // points per scatter work item of the transform being
                         // enqueued: work_chunk(chunk_max, Np)
  int scatter = DSHT_SCATTER_AUTO;       // configured scatter engine of FP32 plans
  int scatter_now = DSHT_SCATTER_MATRIX;  // engine of the transform being enqueued
  int contract = DSHT_CONTRACT_MATRIX;  // engine of batched FP32 contractions
  // contraction layout
  int ngroups = 0;
  GroupDesc* groups = nullptr;  // device [ngroups]
  // the matrix-core contraction's work: pairs of consecutive groups of one
  // (m, P) (nrows up to 128), in XCD order; device, after groups in the same
  // allocation
  int nmgroups = 0;
  GroupDesc* mgroups = nullptr;
  int64_t* mp_off = nullptr;    // device [2*Nl]
  int64_t tab_vec4 = 0;         // entries (row, node pair) of one table copy
  // Y-only tables {Y(x_2np), Y(x_2np+1)} per entry: FP64 plans tab64y, FP32
  // plans tab32y.  tab64: FP64 {Y, Y'} copy of caller-provided tables (csht
  // do_transform), contracted with the caller's Y' as given.
  double* tab64y = nullptr;
  float* tab32y = nullptr;
  // FP32 plans without caller tables hold no Y table: the contraction
  // regenerates Y by the Legendre recurrence (k_contract_otf), from otf_rc:
  // per m at otf_offset(m, Nl), (Nl - m + 32) FP64 pairs {A, B}
  bool otf = false;
  double* otf_rc = nullptr;
  double* tab64 = nullptr;
  bool caller_tables = false;
  int64_t table_bytes = 0;
  // spreading scatter (DSHT_SCATTER_SPREAD): grid of n = 2^spread_logn >= 4 Nl
  // points, kernel parameter beta, tw[k] = e^{-2 pi i k / n} (k < n), and
  // 1/psihat(m) (m < Nl); built for every FP32 plan
  int spread_logn = 0;
  double spread_beta = 0;
  float2* spread_tw = nullptr;
  float* spread_ipsi = nullptr;
  // polar rotation constants, per l at offset l(l+1)/2: d^l_{l,m}(pi/2) as
  // mantissa * 2^exp for m = 0..l, A(m')/B(m') and 1/B(m') of the recursion
  double* rot_top = nullptr;
  int* rot_exp = nullptr;
  double* rot_AB = nullptr;
  double* rot_iB = nullptr;
  // Owner of the device arrays above (groups .. rot_iB); shared by the clones
  // of a plan (dsht_plan_clone), which free them with the last of them.
  std::shared_ptr<TableOwner> table_hold;
  Workspace ws;
  // stage timing (dsht_plan_set_timing): a pool of event sets reused round
  // robin; a set is harvested before reuse and by dsht_plan_get_timing
  static constexpr int kTimingSets = 64;
  bool timing = false;
  std::vector<hipEvent_t> events;     // kTimingSets * (DSHT_NSTAGES + 1)
  std::vector<char> pending;          // per set: recorded, not yet harvested
  int next_set = 0;
  double stage_ms[DSHT_NSTAGES] = {0, 0, 0, 0, 0};
  int64_t timed = 0;
};

// Offset of order m's recurrence coefficients in Plan::otf_rc (pairs), and
// the table's length: entries k = 0 .. Nl - m + 31 for each m.
__host__ __device__ inline int64_t otf_offset(int m, int Nl) {
  return (int64_t)m * (Nl + 32) - (int64_t)m * (m - 1) / 2;
}
// Host-side layout computation shared by plan creation and estimates.
void build_layout(int Nl, int Npair, std::vector<GroupDesc>& groups,
                  std::vector<int64_t>& mp_off, int64_t& tab_vec4);
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X (one L2 each)
std::vector<GroupDesc> xcd_schedule(const std::vector<GroupDesc>& groups, int nxcd);
int64_t workspace_bytes(int Nl, int Nx, int precision, int64_t Np, int chunk);

// Kernel launchers (one per source file).
// out: 0 Y-only FP64 (dst double*), 1 Y-only FP32 (dst float*), 2 FP32 {Y, Y'}
// (dst float*), 3 healpy rows (dst = Yv, hpYd = Yd)
int launch_tables(const Plan& pl, int out, void* dst, double* hpYd, hipStream_t s);
int launch_relayout(const Plan& pl, const double* hpYv, const double* hpYd, hipStream_t s);
// healpy rows row0 .. row0 + nrows - 1 of a Y table (`slab`, device, FP64)
// into the plan's Y-only table
int launch_relayout_y(const Plan& pl, int64_t row0, int64_t nrows, const double* slab,
                      hipStream_t s);
// binning + counting sort; wt NULL = unit weights; flags: the set's 4 status ints
int launch_points(Plan& pl, int64_t Np, const double* theta, const double* phi, const double* wt,
                  int mode, int* flags, hipStream_t s);
// the group's point sets sorted into one record array (segment arrays and
// flags per set: set s at s (nseg + 2) and flags + 4 s; run counts per
// global chunk, run_base over all the group's chunks)
int launch_points_many(Plan& pl, const ManySets& ms, int mode, int* flags, unsigned* wmax,
                       hipStream_t s);
// the matrix-core scatter of the group's chunks; runs keyed by
// set (nseg + 2) + segment in run_seg
int launch_scatter_mfma_many(Plan& pl, const ManySets& ms, const unsigned* wmax, hipStream_t s);
// the fold of the group's sets into rhs slots set0 .. set0 + ms.nsets - 1 of nsets
int launch_fold_many(Plan& pl, const ManySets& ms, int set0, int nsets, hipStream_t s);
// ws.wsort[pos] = w[rec[pos].idx] for the sorted valid points
int launch_gather_weights(Plan& pl, int64_t Np, const double* w, hipStream_t s);
// wsort NULL: the records' own weights
int launch_scatter(Plan& pl, int64_t Np, const double* wsort, hipStream_t s);
// spreading scatter; wmode 0: the records' own weights, 1: w[sorted
// position], 2: w[rec.idx] (a weight set in the caller's input order)
constexpr int kSpreadW = 8;        // kernel taps (beta = 2.3 W)
constexpr int kSpreadMaxNl = 2048;  // grids of n <= 8192 (two 64-bit grids in LDS)
// part: where the run partials go (NULL: ws.part); wmax: max |w| of the
// weights already reduced (NULL: reduced here into ws.wmax when wmode != 0)
int launch_scatter_spread(Plan& pl, int64_t Np, const double* w, int wmode, hipStream_t s,
                          float* part = nullptr, const unsigned* wmax = nullptr);
// the spreading scatter serves FP32 plans up to kSpreadMaxNl; above it the
// matrix engine takes its place
inline bool uses_spread(const Plan& pl) {
  return pl.precision == DSHT_FP32 && pl.scatter_now == DSHT_SCATTER_SPREAD &&
         pl.Nl <= kSpreadMaxNl;
}
// The node variant of the spreading scatter (scatter and fold in one
// launch, one workgroup per node pair) serves grids of at least
// kNodeMinPairs node pairs (fewer leave the device short of work) and at
// most 4096 cells (Nl <= 1024).
constexpr int kNodeMinPairs = 512;
inline bool node_spread_fits(const Plan& pl) {
  return pl.Npair >= kNodeMinPairs && pl.spread_logn <= 12;
}
// DSHT_SCATTER_AUTO: the spreading scatter costs one transform of four
// grids per run (per node and branch in the node variant) and a fixed
// number of grid updates per point, the matrix engine work proportional to
// points x orders, so the spreading scatter is taken where a transform has
// at least this many points per segment on average.  Measured break-even
// (whole transforms, uniform points, Nx = 2 Nl): the node variant 25-50
// points per segment for Nl 512-1024, the run-based one (and the fold
// after it) about 250 at Nl 256.  Weight sets on shared points are
// scattered together on the matrix engine (its operands built once for all
// sets).  The run-based spreading takes them one by one: 0.14 against ~0.25
// ms a set at 2e6 points, Nl = 768 (326 points per segment), 0.78 against
// 0.44 ms at 1e7 points, Nl = 1000 (1221): break-even near 600.  The node
// variant takes them all in one launch: break-even (whole batches, uniform
// points, Nl 512-1000) near 100 points per segment for 16 sets and 250-350
// for 64, about 5 per set (fewer than 4 sets: the plain contraction after
// each set's spreading, 128).
constexpr int kSpreadMinPerSeg = 256;
constexpr int kNodeSpreadMinPerSeg = 64;
constexpr int kSetsSpreadMinPerSeg = 640;
inline int node_sets_min_per_seg(int nsets) { return nsets < 4 ? 128 : 5 * nsets + 16; }
enum PickFor { kPickPlain, kPickPointSet, kPickWeightSets };
inline int pick_scatter(const Plan& pl, int64_t Np, PickFor use = kPickPointSet, int nsets = 1) {
  if (pl.scatter != DSHT_SCATTER_AUTO) return pl.scatter;
  const bool node = node_spread_fits(pl);
  const int per = use == kPickWeightSets ? (node ? node_sets_min_per_seg(nsets) : kSetsSpreadMinPerSeg)
                  : use == kPickPlain && node ? kNodeSpreadMinPerSeg
                                              : kSpreadMinPerSeg;
  return pl.Nl <= kSpreadMaxNl && Np >= (int64_t)per * pl.nseg ? DSHT_SCATTER_SPREAD
                                                               : DSHT_SCATTER_MATRIX;
}
// Work items of the spreading scatter hold up to kSpreadChunk points unless a
// chunk is configured: each run (a stretch of one segment inside an item)
// costs one transform of its grids, so longer items mean fewer runs.
constexpr int kSpreadChunk = 8192;
inline int chunk_cap(const Plan& pl) {
  return uses_spread(pl) && !pl.chunk_explicit ? std::max(pl.chunk_max, kSpreadChunk)
                                               : pl.chunk_max;
}
// *out (NULL: ws.wmax) = max |w[i]| over the finite entries of w[0..Np)
// (float bits)
int launch_weight_max(Plan& pl, int64_t Np, const double* w, hipStream_t s,
                      unsigned* out = nullptr);
// p[0 .. n) = 0 and, if q, q[0 .. n) = 0 (a kernel: capturable)
int launch_clear_u32(unsigned* p, unsigned* q, int n, hipStream_t s);
// gather: w holds a weight set in the caller's input order (read through the
// records' idx); otherwise w is wsort as for launch_scatter
int launch_scatter_mfma(Plan& pl, int64_t Np, const double* w, bool gather, hipStream_t s);
// the plan's scatter gathers weight sets itself (no launch_gather_weights)
inline bool scatter_gathers(const Plan& pl) {
  return pl.precision == DSHT_FP32 &&
         (pl.scatter_now == DSHT_SCATTER_MATRIX || pl.scatter_now == DSHT_SCATTER_SPREAD);
}
// Points per scatter work item of a transform of Np points: the configured
// chunk, or less for a small transform, so that it still spreads over about
// kChunkTarget work items (at least kChunkMin points each, a multiple of 64).
constexpr int kChunkMin = 64;
constexpr int kChunkTarget = 1024;
inline int work_chunk(int chunk_max, int64_t Np) {
  int64_t c = (Np + kChunkTarget - 1) / kChunkTarget;
  c = (c + 63) / 64 * 64;
  return (int)std::max<int64_t>(std::min<int64_t>(kChunkMin, chunk_max),
                                std::min<int64_t>(c, chunk_max));
}
constexpr int kMfmaMinSets = 4;  // FP32 plans: matrix-core contraction from this many sets
constexpr int kMfmaMinPairs = 16;  // and node pairs (one chunk of its K dimension)
constexpr int kMaxSets = 64;     // sets per batched transform
// the matrix-core contraction handles a transform of nsets sets on this plan
inline bool use_mfma_contract(const Plan& pl, int nsets) {
  return pl.precision == DSHT_FP32 && !pl.caller_tables && nsets >= kMfmaMinSets &&
         pl.contract == DSHT_CONTRACT_MATRIX && pl.Npair >= kMfmaMinPairs;
}
// The node variant for transforms that write the plain coefficient layout
// (no matrix-core contraction after them).  rhs: the plain layout
// records (plain_rec); w, wmode as launch_scatter_spread.
inline bool uses_node_spread(const Plan& pl, int nsets) {
  return uses_spread(pl) && !use_mfma_contract(pl, nsets) && node_spread_fits(pl);
}
// The node variant (scatter and fold in one launch) of a single transform:
// the plain records; w, wmode as launch_scatter_spread.
int launch_spread_nodes(Plan& pl, int64_t Np, const double* w, int wmode, hipStream_t s);
// The node variant for sets set0 .. set0 + count - 1 of a batch of ntot sets
// (matrix-core contraction), one launch, into their 16-value row slots, with
// each set's bound into ws.rmax (cleared here): weight_rows, the weight sets of
// launch_place_weight_sets (count of them, maxima in ws.wmax[0 .. count));
// else one set of the records' own weights (count 1, ws.wmax[0]).
int launch_spread_nodes_sets(Plan& pl, int set0, int count, int ntot, bool weight_rows,
                             hipStream_t s);
// The nsets weight sets w + k ldw (k < nsets), FP32, point-major into
// ws.wsf ([input index][nsets]), and max |w| per set into ws.wmax.
int launch_place_weight_sets(Plan& pl, int64_t Np, const double* w, int64_t ldw, int nsets,
                            hipStream_t s);
// the spreading kernel's values at the nq points z (host arrays), as the
// device evaluates them in the scatters (synchronous; at plan creation)
int spread_kernel_values(const float* z, int nq, double beta, float* out);
// weight sets on shared points (DSHT_FP32, matrix scatter engine, from
// kSetsScatterMin sets): one scatter of all sets, set s's run partials at
// part + s * part_set_stride(pl)
constexpr int kSetsScatterMin = 2;
constexpr int kSetsChunk = 16384;  // largest work item of the weight-set scatter
inline bool uses_sets_scatter(const Plan& pl, int nsets) {
  return pl.precision == DSHT_FP32 && pl.scatter_now == DSHT_SCATTER_MATRIX &&
         nsets >= kSetsScatterMin;
}
inline int64_t part_set_stride(const Plan& pl) { return pl.ws.cap_set_runs * (int64_t)pl.Nl * 8; }
// w: nsets rows of ldw weights in the caller's input order
int launch_scatter_sets(Plan& pl, int64_t Np, const double* w, int64_t ldw, int nsets,
                        hipStream_t s);
// node/parity fold of the run partials `part` (NULL: the workspace's first
// slot): slot `set` of nsets for the matrix-core contraction and caller
// tables, else the single plain slot
int launch_fold(Plan& pl, int nsets, int set, hipStream_t s, const void* part = nullptr);
// the fold of weight sets set0 .. set0 + count - 1 (count < 0: all nsets) of
// an nsets batch at once for the matrix-core contraction (FP32), set
// set0 + k's partials at part + k * part_stride
int launch_fold_sets(Plan& pl, int nsets, hipStream_t s, const void* part, int64_t part_stride,
                     int set0 = 0, int count = -1);
// weight-set batches on the spreading scatter spread the sets one by one into
// slots of their own and fold up to this many slots at once
constexpr int kSpreadFoldGroup = 16;
// angular power spectra of nsets alm arrays (dsht_alm2cl)
int launch_alm2cl(int Nl, int nsets, const double* alm, int64_t ld, double* cl, int cross,
                  hipStream_t s);
int launch_threej000_table(int N, double* store, hipStream_t s);
int launch_mode_coupling(int Nl, const double* W, int64_t nW, double* M, hipStream_t s);
int tj_max_n();
int launch_synth_rings(int Nl, const double* alm, int nrings, const double* x, double* F,
                       hipStream_t s);
int synth_max_nl();
int synth_max_rings();
// node ranges of the plain contraction (work items per row group)
#ifndef DSHT_CONTRACT_SPLITS
#define DSHT_CONTRACT_SPLITS 2
#endif
constexpr int kContractSplits = DSHT_CONTRACT_SPLITS;
inline int contract_splits(const Plan& pl) { return pl.Npair / 2 < kContractSplits ? 1 : kContractSplits; }
// plain contraction of the folded set `set` (of nsets) into a_equ / a_pol [set][Nlm][2]
int launch_contract_set(Plan& pl, int nsets, int set, hipStream_t s);
// matrix-core contraction of all nsets slots (FP32 {Y, Y'} table `full32`)
int launch_contract_mfma(Plan& pl, const float* full32, int nsets, hipStream_t s);
int launch_rotate_combine(Plan& pl, int nsets, const int* flags, int flag_stride, double* alm,
                          int* status, hipStream_t s);

}  // namespace dsht
