/*
 * dsht.h -- C ABI of the MI355X direct spherical harmonic transform
 * (libdsht.so, built from directsht_amd/csrc).
 *
 * Two families of entry points:
 *
 *  1. The reference's own native C ABI, kept bit-for-bit in signature and
 *     meaning so a ctypes caller of the reference's csht/sht.py can switch
 *     libraries without code changes (host pointers, caller-owned arrays):
 *       make_table    replaces csht/sht_helper.c:18   (make_table)
 *       do_transform  replaces csht/sht_helper.c:176  (do_transform)
 *
 *  2. A device-resident plan API behind the Python DirectSHT class, which
 *     replaces sht/sht.py:30-284 (DirectSHT.__init__, __call__, basic_sht):
 *     tables stay in HBM between calls, inputs and outputs are device
 *     pointers, work is queued on the caller's HIP stream.
 *
 * All functions return 0 on success and a negative DSHT_E* code on failure;
 * dsht_last_error() returns a message for the calling thread's last failure.
 */
#ifndef DSHT_H
#define DSHT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define DSHT_OK 0
#define DSHT_EINVAL (-1)   /* bad argument                                 */
#define DSHT_EHIP (-2)     /* HIP runtime error                            */
#define DSHT_ENOMEM (-3)   /* device allocation failed                     */

/* ---- status bits written by a transform (dsht_transform *status) ------- */
/* Some point kept by the equatorial branch has theta outside
 * (arccos(xmax), arccos(-xmax)) or theta < 0: the reference raises
 * AssertionError (sht/sht.py:171-172). */
#define DSHT_STATUS_EQU_RANGE 1
/* A polar point is still polar after the 90-degree rotation (possible only
 * when xmax < sin(pi/4)): the reference catches the AssertionError and
 * returns zeros for the whole transform (sht/sht.py:141-149).  The output
 * alm is then all zeros. */
#define DSHT_STATUS_POLAR_RANGE 2

/* ---- precision of the table contraction (tables are always built in FP64) */
#define DSHT_FP64 0
#define DSHT_FP32 1

/* ---- transform modes ---------------------------------------------------- */
/* DirectSHT.__call__: polar/equatorial split, rotation, parity (sht/sht.py:112-150) */
#define DSHT_MODE_CALL 0
/* DirectSHT.basic_sht: no split, every point must be equatorial (sht/sht.py:153-283) */
#define DSHT_MODE_BASIC 1

/* ========================================================================= */
/* 1. Reference C ABI (csht)                                                 */
/* ========================================================================= */

/* Fill caller-owned host arrays Yv, Yd (each Nl(Nl+1)/2 * Nx doubles, healpy
 * row order, row length Nx) with Y_lm(x_j, 0) and dY_lm/dx at
 * x_j = j*xmax/(Nx-1).  Built by the FP64 recurrence on the GPU.  Any
 * Nl >= 2 (memory permitting), 2 <= Nx <= 2^20, 0 < xmax < 1.
 * Replaces csht/sht_helper.c:18-116.  Returns 0 on success, 1 for arguments
 * outside these ranges or a failed device allocation or copy, with the
 * reason in dsht_last_error() (the caller's arrays are then left as they
 * were; the reference returns 0 unconditionally). */
int make_table(int Nl, int Nx, double xmax, double Yv[], double Yd[]);

/* Direct SHT of Np points with cos(theta) = cost[i] (all of one sign,
 * |cost| < xmax), phi[i], weight wt[i], using the caller's tables Yv, Yd (as
 * filled by make_table).  Writes carr[idx] = sum Y cos(m phi) and
 * sarr[idx] = sum Y sin(m phi) for every healpy index idx (alm = carr -
 * i*sarr); the caller applies (-1)^(l-m) for cost < 0, as csht/sht.py:101-121
 * does.  The reference requires cost sorted by |cost|; this implementation
 * sorts on the device and accepts any order.  The tables are uploaded on every
 * call (the plan API keeps them resident).  Replaces csht/sht_helper.c:176-286.
 * Band limit 2 <= Nl <= 4096 (the transform kernels'), 2 <= Nx <= 2^20.
 * Returns 0, or 1 for arguments outside these ranges or failed device work,
 * with the reason in dsht_last_error() and carr, sarr untouched (the
 * reference returns 1 on allocation failure). */
int do_transform(int Nl, int Nx, double xmax, double Yv[], double Yd[], int Np,
                 double cost[], double phi[], double wt[], double carr[], double sarr[]);

/* ========================================================================= */
/* 2. Device-resident plan API                                               */
/* ========================================================================= */

typedef struct dsht_plan dsht_plan;

/* Build the FP64 Ylm/dYlm spline tables of DirectSHT(Nl, Nx, xmax) on the
 * current device (sht/sht.py:32-56) in the layout the contraction streams,
 * keeping an FP64 copy (precision DSHT_FP64) or an FP32 copy (DSHT_FP32).
 * Requires 2 <= Nl <= 4096, 2 <= Nx <= 2^20, 0 < xmax < 1. */
int dsht_plan_create(int Nl, int Nx, double xmax, int precision, dsht_plan** plan);
int dsht_plan_destroy(dsht_plan* plan);

/* dsht_plan_create with the Y table supplied instead of built: Yv is a host
 * array in healpy row order [Nl(Nl+1)/2][Nx], FP64, as dsht_plan_get_tables
 * writes it (the DirectSHT.Yv attribute, sht/sht.py:56).  No counterpart in
 * the reference, which recomputes its tables per instance (sht/sht.py:50-56):
 * it serves an on-disk table cache.  The plan's own table is Yv converted to
 * the contraction layout (rounded to FP32 for DSHT_FP32), so a plan from the
 * tables of an identical plan transforms bit for bit as that plan does. */
int dsht_plan_create_from_tables(int Nl, int Nx, double xmax, int precision, const double* Yv,
                                 dsht_plan** plan);

/* A second plan over the same resident tables, with a workspace of its own,
 * so transforms at the same (Nl, Nx) can run concurrently on different
 * streams (the reference runs one transform at a time; this extends the plan
 * API for servers and batched callers).  The tables are freed with the last
 * plan that shares them. */
int dsht_plan_clone(const dsht_plan* src, dsht_plan** clone);

/* Grow the plan's workspace now to hold a transform of Np points (per set)
 * with up to nsets sets (1 for dsht_transform), so that later transforms
 * within those sizes allocate nothing and never synchronize (what HIP graph
 * capture and latency-bound callers need).  Growth waits for the device.
 * On an FP32 plan with nsets >= 4 and the matrix-core contraction this also
 * builds the {Y, Y'} table that contraction reads (otherwise built by the
 * first such transform). */
int dsht_plan_reserve(dsht_plan* plan, int64_t Np, int nsets);

/* dsht_plan_reserve for some kinds of transform only (no counterpart in the
 * reference; it spares batched callers the workspace of the kinds they never
 * run): DSHT_RESERVE_POINT_SETS for dsht_transform and dsht_transform_many
 * (up to nsets point sets of up to Np points each), DSHT_RESERVE_WEIGHT_SETS
 * for dsht_transform_sets (up to nsets weight sets on Np shared points), or
 * both (= dsht_plan_reserve). */
#define DSHT_RESERVE_POINT_SETS 1
#define DSHT_RESERVE_WEIGHT_SETS 2
int dsht_plan_reserve_for(dsht_plan* plan, int64_t Np, int nsets, int kinds);

/* Bytes of HBM the plan holds: its tables (shared with its clones, see
 * dsht_plan_clone; the Y-only table, plus the FP32 {Y, Y'} table once a
 * batched matrix-core contraction has built it) + its own current
 * workspace. */
int dsht_plan_footprint(const dsht_plan* plan, int64_t* bytes);

/* HBM a plan would need for its tables and for a transform of Np points
 * (what a server admits against).  table_bytes covers the Y-only table the
 * single-set contraction streams; a batched FP32 transform with the
 * matrix-core contraction adds a {Y, Y'} table of twice its size. */
int dsht_plan_estimate(int Nl, int Nx, int precision, int64_t Np, int64_t* table_bytes,
                       int64_t* work_bytes);

/* Copy the tables to caller-owned host arrays in healpy row order (the
 * DirectSHT.Yv / .Yd attributes, sht/sht.py:56), FP64.  Either may be NULL. */
int dsht_plan_get_tables(dsht_plan* plan, double* Yv, double* Yd);

/* Transform Np points.  theta, phi, wt: device arrays of Np doubles.
 * alm: device array of Nl(Nl+1)/2 complex doubles (interleaved re, im),
 * healpy order, overwritten.  status: device int written with DSHT_STATUS_*
 * bits (may be NULL).  mode: DSHT_MODE_CALL or DSHT_MODE_BASIC.  Work is
 * enqueued on `stream` (a hipStream_t; 0 = default stream) and the call
 * returns without synchronizing (unless the workspace has to grow; see
 * dsht_plan_reserve).  Concurrent transforms on one plan must be serialized
 * by the caller (they share the plan's workspace); clones do not share it. */
int dsht_transform(dsht_plan* plan, int64_t Np, const double* theta, const double* phi,
                   const double* wt, double* alm, int* status, int mode, void* stream);

/* dsht_transform without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during a
 * transform (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results as
 * dsht_transform; the plan's tables are used as they are, and the transform's
 * workspace is carved from a fixed 64 MiB device buffer inside libdsht.so
 * (one per device, held until the process ends) instead of the plan's
 * workspace, which is neither used, grown nor freed.  This covers Nl <= 256:
 * e.g. up to about 2.3e5 points at Nl=256, Nx=512 (FP32) or 8.4e5 at Nl=64,
 * Nx=128.  Larger transforms return DSHT_EINVAL with the bytes they would
 * need in dsht_last_error(), and leave alm untouched.  Every call on a device
 * shares that device's buffer: calls whose work overlaps on one device (other
 * streams, other threads, other plans) race, so serialize them.  The buffer
 * keeps the last call's contents; the kernels clear what must start at zero. */
int dsht_transform_noalloc(dsht_plan* plan, int64_t Np, const double* theta, const double* phi,
                           const double* wt, double* alm, int* status, int mode, void* stream);

/* Batched transforms (no counterpart in the reference, which transforms one
 * weight vector of one point set per call; these share the resident tables
 * and read them once for the whole batch).  At most 64 sets per call.
 *
 * Many weight sets on shared points: alm[s] = transform of (theta, phi,
 * wt + s * ldw) for s < nsets (ldw >= Np).  The points are binned and sorted
 * once.  alm: device array [nsets][Nlm] complex doubles; status: device
 * [nsets] ints (may be NULL). */
int dsht_transform_sets(dsht_plan* plan, int64_t Np, const double* theta, const double* phi,
                        int nsets, const double* wt, int64_t ldw, double* alm, int* status,
                        int mode, void* stream);

/* Many independent point sets at the plan's (Nl, Nx): set s has Np[s] points
 * theta[s], phi[s], wt[s] (host arrays of device pointers).  alm, status as
 * for dsht_transform_sets.  The workspace is reused set after set; the table
 * contraction and the polar rotation run once for all sets. */
int dsht_transform_many(dsht_plan* plan, int nsets, const int64_t* Np, const double* const* theta,
                        const double* const* phi, const double* const* wt, double* alm,
                        int* status, int mode, void* stream);

/* Tuning options (dsht_plan_configure).  Results do not depend on them beyond
 * rounding.
 *   DSHT_OPT_SCATTER  scatter engine of DSHT_FP32 plans:
 *                     DSHT_SCATTER_AUTO (default) picks per transform: the
 *                     spreading engine where Nl <= 2048 and the points
 *                     average at least k per segment of the (Nx - 1)-bin grid
 *                     (4 (Nx - 1) segments): k = 64 for a single transform
 *                     with 512 or more node pairs and Nl <= 1024, 256 for
 *                     other transforms and the dense sets of a
 *                     dsht_transform_many batch, 640 for dsht_transform_sets
 *                     batches; else the matrix engine;
 *                     DSHT_SCATTER_MATRIX fp16x3 split-precision MFMA GEMM;
 *                     DSHT_SCATTER_VECTOR FP32 VALU recurrence;
 *                     DSHT_SCATTER_SPREAD: the sums as type-1 non-uniform
 *                     FFTs (kernel spreading onto a 4x oversampled phi grid
 *                     in 64-bit fixed point, FFT, deconvolution; per node
 *                     with the fold included, or per run; ~1e-7 relative;
 *                     Nl <= 2048, the matrix engine above).
 *                     DSHT_FP64 plans always use the FP64 vector scatter.
 *   DSHT_OPT_CHUNK    sorted points per scatter work item (64..65536; default
 *                     2048, and 8192 for the spreading engine).  A transform of fewer than 1024 x chunk points
 *                     uses smaller items (about Np / 1024, at least 64) so
 *                     that it still spreads over the whole device.
 *   DSHT_OPT_CONTRACT table contraction of batched DSHT_FP32 transforms with
 *                     4 or more sets: DSHT_CONTRACT_MATRIX (default) fp16x3
 *                     matrix-core GEMM reading the tables once for all sets,
 *                     or DSHT_CONTRACT_VECTOR FP32 GEMV once per set.      */
#define DSHT_OPT_SCATTER 1
#define DSHT_OPT_CHUNK 2
#define DSHT_OPT_CONTRACT 3
#define DSHT_SCATTER_MATRIX 0
#define DSHT_SCATTER_VECTOR 1
#define DSHT_SCATTER_SPREAD 2
#define DSHT_SCATTER_AUTO 3
#define DSHT_CONTRACT_MATRIX 0
#define DSHT_CONTRACT_VECTOR 1
int dsht_plan_configure(dsht_plan* plan, int option, int value);
/* The scatter engine (DSHT_SCATTER_*) a single-set transform of Np points
 * takes on this plan: the configured one, DSHT_SCATTER_AUTO resolved by the
 * rule above, DSHT_SCATTER_MATRIX for DSHT_SCATTER_SPREAD above Nl = 2048,
 * and DSHT_SCATTER_VECTOR for DSHT_FP64 plans.  No counterpart in the
 * reference (its numpy path has one scatter); for callers that report or
 * tune per engine. */
int dsht_plan_engine(const dsht_plan* plan, int64_t Np, int* engine);

/* Stage timing.  While enabled, every dsht_transform records HIP events on
 * its stream around its five stages: 0 point binning + counting sort,
 * 1 scatter, 2 node/parity fold, 3 table contraction, 4 polar rotation +
 * combine.  dsht_plan_get_timing waits for the recorded transforms and
 * returns the summed milliseconds per stage and the number of transforms
 * summed.  Enabling resets the sums. */
#define DSHT_NSTAGES 5
int dsht_plan_set_timing(dsht_plan* plan, int enable);
int dsht_plan_get_timing(dsht_plan* plan, double stage_ms[DSHT_NSTAGES], int64_t* ntransforms);

/* Diagnostic: copy the node coefficients of the plan's last single-set
 * transform (the right-hand side of the contraction, after the scatter and
 * fold) to `out`: Nl * 2 * Npair * 24 values, Npair = 2 ceil(Nx/4) node pairs
 * (nodes past Nx are zero padding), float for DSHT_FP32 plans,
 * double for DSHT_FP64, layout [node pair][branch][m][parity P]
 * [kind*4 + part*2 + node] with branch 0 equatorial, 1 polar, kind c (the Y
 * coefficients c_P), f (x d_P / (1 - x^2)) and g (d_{1-P} / (1 - x^2)), d the
 * Y' coefficients, part re, im, node = even, odd node of the pair.  Waits
 * for the device. */
int dsht_plan_debug_rhs(dsht_plan* plan, void* out, int64_t bytes);

/* Angular power spectra of nsets alm arrays on the device (healpy order,
 * complex doubles; set s at alm + 2*s*ld doubles, ld >= Nl(Nl+1)/2), written
 * to the device array cl, enqueued on `stream`, nothing synchronized.
 *   cross == 0: cl[s*Nl + l], the auto-spectrum of set s,
 *     C_l = [(Re a_l0)^2 + 2 sum_{m>0} |a_lm|^2] / (2l+1)
 *     -- DirectSHT.alm2cl, sht/sht.py:95-109;
 *   cross != 0: cl[p*Nl + l] for the nsets(nsets+1)/2 pairs p = (s, t),
 *     s <= t, in row order ((0,0), (0,1), .., (0,S-1), (1,1), ..),
 *     C^{st}_l = [Re a^s_l0 Re a^t_l0 + 2 sum_{m>0} Re(a^s_lm conj(a^t_lm))] / (2l+1)
 *     -- healpy alm2cl(alm1, alm2) as its callers use it
 *     (notebooks/analyzing_mocks.ipynb), m = 0 real parts as above.
 * 1 <= nsets <= 64.  No counterpart in csht. */
int dsht_alm2cl(int Nl, int nsets, const double* alm, int64_t ld, double* cl, int cross,
                void* stream);

/* Wigner 3j symbols (j1 j2 j3; 0 0 0) for N > j1 >= j2 >= j3 >= 0, written
 * to the device array store of N(N+1)(N+2)/6 doubles at
 * j1(j1+1)(j1+2)/6 + j2(j2+1)/2 + j3 (zero for odd j1+j2+j3 or j3 < j1-j2),
 * enqueued on `stream`.  1 <= N <= 4096.  The first call on a device uploads a
 * 96 KB coefficient table (synchronously).
 * Replaces make_table of csht/threej000_helper.c:62 (a separate library in
 * the reference, whose symbol name clashes with csht's make_table) and
 * Wigner3j.get_3js, sht/threej000.py:107-116. */
int dsht_threej000_table(int N, double* store, void* stream);

/* Mode-coupling matrix of a window function, on the device:
 *   M[l1*Nl + l2] = (2 l2 + 1)/(4 pi) sum_{l3} (2 l3 + 1) W[l3] (l1 l2 l3; 0 0 0)^2
 * for 0 <= l1, l2 < Nl, with W (device, nW doubles) taken as zero from nW on
 * (the reference zero-pads it to 2Nl-1).  Enqueued on `stream`.
 * 1 <= Nl <= 4096, nW >= 0.  First call on a device: as dsht_threej000_table.
 * Replaces MaskDeconvolution.get_M, sht/mask_deconvolution.py:171-192. */
int dsht_mode_coupling(int Nl, const double* W, int64_t nW, double* M, void* stream);

/* Harmonic synthesis on iso-latitude rings, on the device:
 *   F[r*Nl + m] = sum_{l=m}^{Nl-1} a_lm Ybar_lm(x[r])     (complex doubles)
 * for alm in healpy order (Nl(Nl+1)/2 complex doubles), ring cosines x
 * (nrings doubles in [-1, 1]) and Ybar the fully normalized associated
 * Legendre function with the Condon-Shortley phase; a real field on ring r
 * is then Re F[r][0] + 2 Re sum_{m>0} F[r][m] e^{i m phi}.  Enqueued on
 * `stream`.  1 <= Nl <= 2048, 0 <= nrings <= 67107840.
 * Stands in for healpy synfast in the reference's lognormal mocks
 * (sht/mocks.py:60); no counterpart in csht. */
int dsht_synthesis_rings(int Nl, const double* alm, int nrings, const double* x, double* F,
                         void* stream);

/* Message describing the calling thread's last error ("" if none). */
const char* dsht_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DSHT_H */
