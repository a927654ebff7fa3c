// Plan: resident tables + per-transform workspace of one DirectSHT(Nl, Nx, xmax).
#pragma once

#include <memory>
#include <vector>

#include "common.h"

namespace dsht {

struct Workspace {
  int64_t cap_points = 0;  // capacity of the sorted-point arrays
  int64_t cap_chunks = 0;  // capacity of per-chunk arrays
  int64_t cap_runs = 0;    // capacity of the run partials
  int64_t cap_tiles = 0;   // capacity of the tile histogram matrix (rows)
  int cap_sets = 0;        // sets of a batched transform the per-set arrays hold
  bool has_wsort = false;  // wsort allocated (weight-set transforms)
  int cur_chunks = 0;      // scatter chunks of the transform being enqueued
  PointRec* rec = nullptr;    // [cap_points] sorted point records
  int* run_count = nullptr;   // [cap_chunks + 1]: runs per chunk
  int* run_base = nullptr;    // [cap_chunks + 1]: exclusive scan of run_count
  int* run_seg = nullptr;     // [cap_runs]: segment of each run
  void* part = nullptr;       // [cap_runs][4][Nl][2]: per-run v_k(m), k = 0..3, complex;
                              // double (FP64 scatter) or float (FP32 scatter)
  int* seg_count = nullptr;   // [nseg + 2]
  int* seg_start = nullptr;   // [nseg + 2]
  int* seg_cursor = nullptr;  // [nseg + 2]
  int* seg_run_lo = nullptr;  // [nseg + 2]: first run of each segment (0 if empty)
  int* seg_run_hi = nullptr;  // [nseg + 2]: one past its last run (0 if empty)
  int* ne_rank = nullptr;     // [nseg + 1]: non-empty segments before s
  int* thist = nullptr;       // [cap_tiles][nseg + 1]: tile histograms, then offsets
  void* rhs = nullptr;        // [Nl][2][Npair][cap_sets][16] float or double
  unsigned* rmax = nullptr;   // [cap_sets]: max(|c|, |d/dx|) per set (float bits)
  double* wsort = nullptr;    // [cap_points]: one weight set in sorted order
  double* a_equ = nullptr;    // [cap_sets][Nlm][2]
  double* a_pol = nullptr;    // [cap_sets][Nlm][2]
  int* flags = nullptr;       // [cap_sets][4]: status bits, n_polar, unused, unused
  int64_t bytes = 0;
};

// Frees a plan's table arrays (on the plan's device) when destroyed.
struct TableOwner {
  int device;
  std::vector<void*> ptrs;
  TableOwner(int dev, std::vector<void*> p) : device(dev), ptrs(std::move(p)) {}
  TableOwner(const TableOwner&) = delete;
  TableOwner& operator=(const TableOwner&) = delete;
  ~TableOwner() {
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != device) (void)hipSetDevice(device);
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
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
  int chunk = 2048;  // sorted points per scatter work item
  int scatter = DSHT_SCATTER_MATRIX;  // engine of FP32 plans
  int contract = DSHT_CONTRACT_MATRIX;  // engine of batched FP32 contractions
  // contraction layout
  int ngroups = 0;
  GroupDesc* groups = nullptr;  // device [ngroups]
  int64_t* mp_off = nullptr;    // device [2*Nl]
  int64_t tab_vec4 = 0;         // 4-vectors in one table copy
  double* tab64 = nullptr;
  float* tab32 = nullptr;
  int64_t table_bytes = 0;
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

// Host-side layout computation shared by plan creation and estimates.
void build_layout(int Nl, int Npair, std::vector<GroupDesc>& groups,
                  std::vector<int64_t>& mp_off, int64_t& tab_vec4);
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X (one L2 each)
std::vector<GroupDesc> xcd_schedule(const std::vector<GroupDesc>& groups, int nxcd);
int64_t workspace_bytes(int Nl, int Nx, int precision, int64_t Np, int chunk);

// Kernel launchers (one per source file).
int launch_tables(const Plan& pl, bool w64, bool w32, double* hpYv, double* hpYd, hipStream_t s);
int launch_relayout(const Plan& pl, const double* hpYv, const double* hpYd, hipStream_t s);
// binning + counting sort; wt NULL = unit weights; flags: the set's 4 status ints
int launch_points(Plan& pl, int64_t Np, const double* theta, const double* phi, const double* wt,
                  int mode, int* flags, hipStream_t s);
// ws.wsort[pos] = w[rec[pos].idx] for the sorted valid points
int launch_gather_weights(Plan& pl, int64_t Np, const double* w, hipStream_t s);
// wsort NULL: the records' own weights
int launch_scatter(Plan& pl, int64_t Np, const double* wsort, hipStream_t s);
int launch_scatter_mfma(Plan& pl, int64_t Np, const double* wsort, hipStream_t s);
// node/parity fold of the last scatter into slot `set` of nsets
int launch_fold(Plan& pl, int nsets, int set, hipStream_t s);
// contraction of all nsets slots into a_equ / a_pol [set][Nlm][2]
constexpr int kMfmaMinSets = 4;  // FP32 plans: matrix-core contraction from this many sets
constexpr int kMaxSets = 64;     // sets per batched transform
int launch_contract(Plan& pl, int nsets, hipStream_t s);
int launch_contract_mfma(Plan& pl, int nsets, hipStream_t s);
int launch_rotate_combine(Plan& pl, int nsets, const int* flags, int flag_stride, double* alm,
                          int* status, hipStream_t s);

}  // namespace dsht
