// Stage (c): contraction of the resident tables with the node coefficients,
//   a_lm = sum_n Y_lm(x_n) c_P(m, n) + Y'_lm(x_n) d_P(m, n),  P = (l-m) & 1,
// for both branches (equatorial, polar) at once -- the per-(l,m) dot products
// of sht/interp_funcs_py.py:46-60 / sht/interp_funcs_jax.py:31-48 and
// csht/sht_helper.c:255-274, as one HBM-streaming GEMV.
//
// The plan's own tables hold Y only; the Y' term comes from the derivative
// relation (the fold's coefficients f, g):
//   sum_n Y'_l d_P = G_l sum_n Y_{l-1} g - l sum_n Y_l f_P,
// so each row l accumulates S_c = sum Y_l c_P, S_f = sum Y_l f_P and
// S_g = sum Y_l g (g for the row l+1 of the other parity); k_add_der then
// forms a_lm = S_c - l S_f + G_l S_g(l-1).  Half the table bytes of a
// {Y, Y'} table.
//
// One wavefront per row group (up to 64 rows l of one (m, P)), launched in an
// XCD-aware order (xcd_schedule in api.hip); lane r owns row
// l = m + P + 2(i0 + r) and walks the node pairs.  Each step loads one pair
// {Y(x_2np), Y(x_2np+1)} per lane (a contiguous 512-byte (FP32) / 1 KB (FP64)
// block per wavefront) and multiplies it with the wave-uniform coefficients of
// that pair, which the compiler keeps in scalar registers.  FP32 tables
// accumulate 16 node pairs in FP32 and then in FP64; FP64 tables accumulate
// in FP64.  Results go straight to the healpy index of (l, m).
//
// Caller-provided {Y, Y'} tables (csht do_transform) are contracted as given
// by k_contract64.
#include "common.h"
#include "plan.h"

namespace dsht {

// Table entries are read exactly once per transform: non-temporal loads keep
// the stream from displacing the node coefficients (reused by every row
// group of an (m, P)) in the caches.
__device__ inline double4 stream_load(const double4* p) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_double4(v.x, v.y, v.z, v.w);
}

__device__ inline float4 stream_load(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// Y-only tables are read in couples: {Y(x_4c), Y(x_4c+1), Y(x_4c+2),
// Y(x_4c+3)} of one row = node pairs 2c and 2c+1 (one 16-byte (FP32) /
// 32-byte (FP64) load per lane, a contiguous 1 KB / 2 KB block per wavefront).
template <typename T>
struct Couple;
template <>
struct Couple<float> {
  using type = float4;
  static constexpr int kRing = 8;  // couples in flight per lane = 16 pairs per FP32 partial sum
};
template <>
struct Couple<double> {
  using type = double4;
  static constexpr int kRing = 4;
};

// Table loads run kRing couples ahead of the multiply (a ring of registers),
// so that each wavefront keeps kRing loads in flight; lanes past the group's
// rows read row 0 and discard their sums (no branch anywhere).  Scheduling
// fences keep one node pair's coefficients (24 scalar registers) live at a
// time, plus, for FP32, the next pair's, loaded ahead.
//
// Work item = (row group, range of couples): the couples are split into
// gridDim.y ranges, so that the launch has several times more equal-length
// wavefronts than the GPU holds at once (one wavefront per group would run
// in ~1.4 rounds, the second mostly idle).  Each item writes its partial
// sums part[range][h][8] = {S_c - l S_f (4), S_g (4)}; k_add_der adds the
// ranges in order (deterministic).
template <typename T>
__global__ __launch_bounds__(64) void k_contract_y(const typename Couple<T>::type* __restrict__ tab,
                                                   const T* __restrict__ rhs,
                                                   const GroupDesc* __restrict__ groups, int Npair,
                                                   int Nl, int64_t Nlm, double* __restrict__ part) {
  using T4 = typename Couple<T>::type;
  constexpr int kRing = Couple<T>::kRing;
  const GroupDesc gd = groups[blockIdx.x];
  if (gd.nrows == 0) return;  // padding of the XCD schedule
  const int ncpl = Npair / 2;
  const int nsplit = gridDim.y, split = blockIdx.y;
  const int lo = (int)((int64_t)ncpl * split / nsplit);
  const int hi = (int)((int64_t)ncpl * (split + 1) / nsplit);
  const int r = threadIdx.x;
  const bool valid = r < gd.nrows;
  const T4* t = tab + gd.off / 2 + (valid ? r : 0);  // gd.off is even (Npair is)
  // c[8 kind + 2 (2 br + ri) + node]: [c | f | g][er, ei, pr, pi][node] of
  // node pair np, from its two branch records
  const int64_t stride = gd.nrows;
  T4 ring[kRing];
#pragma unroll
  for (int k = 0; k < kRing; ++k) ring[k] = stream_load(t + (int64_t)min(lo + k, hi - 1) * stride);
  double S[3][4] = {};
  auto coef = [&](int np, T (&c)[24]) {
#pragma unroll
    for (int br = 0; br < 2; ++br) {
      const T* qq = rhs + plain_rec(np, br, gd.m, gd.P, Nl);
#pragma unroll
      for (int i = 0; i < 12; ++i) c[8 * (i >> 2) + 4 * br + (i & 3)] = qq[i];
    }
  };
  auto mul = [&](T (&f)[3][4], T y0, T y1, const T (&c)[24]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int o = 0; o < 4; ++o)
        f[k][o] = fma(y0, c[8 * k + 2 * o], fma(y1, c[8 * k + 2 * o + 1], f[k][o]));
  };
  auto flush = [&](T (&f)[3][4]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int o = 0; o < 4; ++o) S[k][o] += (double)f[k][o];
  };
  // couple c: pairs 2c (coefficients in ca) and 2c+1 (cb)
  constexpr bool kPre = sizeof(T) == 4;  // FP32: next pair's coefficients loaded ahead
  T ca[24], cb[24];
  if constexpr (kPre) coef(2 * lo, ca);
  // (lanes past the group's rows compute on row 0 and never store)
  auto couple = [&](T (&f)[3][4], T4 y, int c) {
    if constexpr (kPre) {
      coef(2 * c + 1, cb);
      __builtin_amdgcn_sched_barrier(0);  // issue the loads before the multiplies
      mul(f, y.x, y.y, ca);
      __builtin_amdgcn_sched_barrier(0);
      coef(min(2 * c + 2, 2 * hi - 1), ca);
      __builtin_amdgcn_sched_barrier(0);
      mul(f, y.z, y.w, cb);
    } else {
      coef(2 * c, ca);
      mul(f, y.x, y.y, ca);
      __builtin_amdgcn_sched_barrier(0);
      coef(2 * c + 1, cb);
      mul(f, y.z, y.w, cb);
    }
  };
  int c0 = lo;
  for (; c0 + kRing <= hi; c0 += kRing) {  // full rounds of the ring
    T f[3][4] = {};
#pragma unroll
    for (int k = 0; k < kRing; ++k) {
      // consume the slot, then refill the same registers kRing couples ahead
      couple(f, ring[k], c0 + k);
      ring[k] = stream_load(t + (int64_t)min(c0 + kRing + k, hi - 1) * stride);
      __builtin_amdgcn_sched_barrier(0);
    }
    flush(f);
  }
  {  // the last hi - c0 < kRing couples are already in the ring
    T f[3][4] = {};
#pragma unroll
    for (int k = 0; k < kRing; ++k)
      if (c0 + k < hi) couple(f, ring[k], c0 + k);
    flush(f);
  }
  if (valid) {
    const int ell = gd.m + gd.P + 2 * (gd.i0 + r);
    const int64_t h = hp_index(ell, gd.m, Nl);
    double4* o = reinterpret_cast<double4*>(part + ((int64_t)split * Nlm + h) * 8);
    o[0] = make_double4(S[0][0] - ell * S[1][0], S[0][1] - ell * S[1][1], S[0][2] - ell * S[1][2],
                        S[0][3] - ell * S[1][3]);
    o[1] = make_double4(S[2][0], S[2][1], S[2][2], S[2][3]);
  }
}

// FP32 plans: the same sums without a table.  Y_lm(x_n) is regenerated by
// the normalized Legendre recurrence in l (FP64, the coefficients of
// otf_coefficients in api.hip, three FP64 operations per (l, m, node)), and
// the products with the node coefficients run on the FP32 matrix cores:
// per row tile (16 rows of one parity) and node quad,
//   D[row][col] += sum_{k<4} Y(row, node 4q + k) B[k][col],
// B = the 12 coefficients c, f, g x (er, ei, pr, pi) of the row's parity
// (columns 12..15 unused), v_mfma_f32_16x16x4f32.  No table is read: the
// kernel streams nothing but the coefficients (once per workgroup).
//
// Workgroup = (order m, range of 512 nodes), 4 wavefronts of 128 nodes,
// each lane two chains (nodes 64 c + lane of its wavefront).  A step
// generates 32 rows l = m + 32 s + t (t even: parity 0, odd: parity 1) into
// the wavefront's LDS tiles, rounded to FP32, in the A-operand order (lane
// k * 16 + i reads row i, nodes 4 q + k for four consecutive quads q as one
// 16-byte read; rows padded to 132 floats and quads XOR-swizzled by k, so
// neither the writes nor the reads conflict in LDS banks), and multiplies
// them out: per wavefront FP32 sums over its 128 nodes, summed over the 4
// wavefronts in FP64 and written as the partial sums of k_contract_y (the
// node ranges are dealt to the gridDim.y workgroup columns, and a column
// with several ranges adds the later ones to its sums: the same threads, in
// order).  76 KB of LDS, so two workgroups share a CU; not synchronized with
// each other, one's recurrence (vector unit) can run beside the other's
// products (matrix cores).  (Double-buffered tiles with the next step's
// recurrence interleaved between the MFMAs in each wavefront, at one
// workgroup per CU, measured 1.5x slower.)
constexpr int kOtfWaves = 4, kOtfChains = 2;
constexpr int kOtfNodes = kOtfWaves * 64 * kOtfChains;  // nodes per workgroup
constexpr int kOtfRow = 132;                            // floats per staged row
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kOtfWaves * 64) void k_contract_otf(
    const double2* __restrict__ rc, const float* __restrict__ rhs, int Nl, int Nx, int Npair,
    double dx, int64_t Nlm, double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) float stage[kOtfWaves][2][16 * kOtfRow];
  __shared__ float dsum[kOtfWaves][2][256];
  const int m = blockIdx.x, split = blockIdx.y;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const double2* __restrict__ rcm = rc + otf_offset(m, Nl);
  const int nranges = (2 * Npair + kOtfNodes - 1) / kOtfNodes;
  const int kk = lane >> 4, col = lane & 15;
  // column col = 4 kind + 2 br + ri of the B operand: branch br's record,
  // value 4 kind + 2 ri + node (columns 12..15 load a valid, unused value)
  const int cbr = col < 12 ? (col & 3) >> 1 : 0;
  const int cofs = col < 12 ? 4 * (col >> 2) + 2 * (col & 1) : 0;
  // where this lane writes node 64 c + lane (k, quad) and reads (row i, k)
  const int wk = lane & 3, wq = lane >> 2;
  const int ri = lane & 15, rk = lane >> 4;
  const int nsteps = (Nl - m + 31) / 32;
  for (int rg = split; rg < nranges; rg += gridDim.y) {
    // chains: Y_{l-2}, Y_{l-1} (ya, yb) at x_n, from Y_mm and Y_{m+1,m}
    double x[kOtfChains], ya[kOtfChains], yb[kOtfChains];
#pragma unroll
    for (int c = 0; c < kOtfChains; ++c) {
      const int n = rg * kOtfNodes + w * 128 + c * 64 + lane;
      const bool ok = n < Nx;  // padding nodes (odd Nx, past Nx) hold Y = 0
      x[c] = ok ? (double)n * dx : 0.0;
      const double ymm = ok ? rcm[0].x * pow(sqrt(1.0 - x[c] * x[c]), (double)m) : 0.0;
      ya[c] = ymm;
      yb[c] = rcm[1].x * x[c] * ymm;
    }
    // B operands: lane (k = lane / 16, col = lane % 16) holds column col of
    // node 4 q + k of this wavefront, per parity and quad (loaded once per range)
    float bop[2][32];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 32; ++q) {
        const int n = rg * kOtfNodes + w * 128 + 4 * q + kk;
        const int np = min(n >> 1, Npair - 1);  // past the nodes: Y = 0 there
        bop[p][q] = rhs[plain_rec(np, cbr, m, p, Nl) + cofs + (n & 1)];
      }
    float* tile = &stage[w][0][0];
    // rows t0 .. t0+3 of step s into the wavefront's tiles
    auto gen4 = [&](int s, int t0) {
#pragma unroll
      for (int dt = 0; dt < 4; ++dt) {
        const int t = t0 + dt;
        const bool start = s == 0 && t < 2;  // rows m, m+1: the start values
        const double2 ab = start ? make_double2(0.0, 0.0) : rcm[32 * s + t];
        float* row = tile + (t & 1) * 16 * kOtfRow + (t >> 1) * kOtfRow + wk * 32;
#pragma unroll
        for (int c = 0; c < kOtfChains; ++c) {
          double y;
          if (start) {
            y = t == 0 ? ya[c] : yb[c];
          } else {
            y = fma(ab.x * x[c], yb[c], -(ab.y * ya[c]));
            ya[c] = yb[c];
            yb[c] = y;
          }
          row[(16 * c + wq) ^ (8 * wk)] = (float)y;
        }
      }
    };
    for (int s = 0; s < nsteps; ++s) {
#pragma unroll
      for (int t0 = 0; t0 < 32; t0 += 4) gen4(s, t0);
      __syncthreads();  // the tiles are written (and the last step's sums read)
      // eight independent accumulators (parity x quad mod 4), so consecutive
      // MFMAs never wait for each other
      f32x4_t acc[2][4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[p][u] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q0 = 0; q0 < 32; q0 += 4) {
        f32x4_t a4[2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
          a4[p] = *reinterpret_cast<const f32x4_t*>(tile + p * 16 * kOtfRow + ri * kOtfRow +
                                                    rk * 32 + (q0 ^ (8 * rk)));
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int p = 0; p < 2; ++p)
            acc[p][u] =
                __builtin_amdgcn_mfma_f32_16x16x4f32(a4[p][u], bop[p][q0 + u], acc[p][u], 0, 0, 0);
      }
      // D of lane (k, col), register r: row 4 k + r, column col, kept at
      // [r][k][col] (the 64 lanes of a store on 64 banks)
      float* ds = &dsum[w][0][0];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          ds[p * 256 + r * 64 + rk * 16 + ri] =
              (acc[p][0][r] + acc[p][1][r]) + (acc[p][2][r] + acc[p][3][r]);
      __syncthreads();  // the wavefronts' sums are in LDS
      if (tid < 128) {  // (parity, row, part o): S_c - l S_f and S_g over the 4 wavefronts
        const int p = tid >> 6, rr = (tid >> 2) & 15, o = tid & 3;
        const int ell = m + 32 * s + 2 * rr + p;
        if (ell < Nl) {
          double sc = 0.0, sf = 0.0, sg = 0.0;
          const int d0 = p * 256 + (rr & 3) * 64 + (rr >> 2) * 16;  // row rr, [r][k][col]
#pragma unroll
          for (int v = 0; v < kOtfWaves; ++v) {
            const float* dv = &dsum[v][0][0];
            sc += (double)dv[d0 + o];
            sf += (double)dv[d0 + 4 + o];
            sg += (double)dv[d0 + 8 + o];
          }
          double* out = part + ((int64_t)split * Nlm + hp_index(ell, m, Nl)) * 8;
          if (rg == split) {
            out[o] = sc - ell * sf;
            out[4 + o] = sg;
          } else {
            out[o] += sc - ell * sf;
            out[4 + o] += sg;
          }
        }
      }
    }
    __syncthreads();  // the next range's tiles and sums overwrite this range's
  }
}

// a_lm = sum over ranges of (S_c - l S_f) + G_l sum over ranges of S_g(l-1, m)
// (l > m), G_l = sqrt((l^2 - m^2)(2l+1)/(2l-1)): the partial sums of
// k_contract_y combined with the rest of the Y' term (thread per (l, m))
__global__ __launch_bounds__(256) void k_add_der(int Nl, int64_t Nlm, int nsplit, int64_t set_off,
                                                 const double* __restrict__ part,
                                                 double* __restrict__ a_equ,
                                                 double* __restrict__ a_pol) {
  const int ell = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (ell < m || ell >= Nl) return;
  const int64_t h = hp_index(ell, m, Nl);
  double4 a = make_double4(0, 0, 0, 0), d = make_double4(0, 0, 0, 0);
  for (int sp = 0; sp < nsplit; ++sp) {
    const double4 v = *reinterpret_cast<const double4*>(part + ((int64_t)sp * Nlm + h) * 8);
    a.x += v.x;
    a.y += v.y;
    a.z += v.z;
    a.w += v.w;
    if (ell > m) {
      const double4 u = *reinterpret_cast<const double4*>(part + ((int64_t)sp * Nlm + h - 1) * 8 + 4);
      d.x += u.x;
      d.y += u.y;
      d.z += u.z;
      d.w += u.w;
    }
  }
  if (ell > m) {
    const double G = sqrt((double)(ell - m) * (double)(ell + m) * (double)(2 * ell + 1) /
                          (double)(2 * ell - 1));
    a.x += G * d.x;
    a.y += G * d.y;
    a.z += G * d.z;
    a.w += G * d.w;
  }
  const int64_t idx = h + set_off;
  a_equ[2 * idx] = a.x;
  a_equ[2 * idx + 1] = a.y;
  a_pol[2 * idx] = a.z;
  a_pol[2 * idx + 1] = a.w;
}

__global__ __launch_bounds__(64) void k_contract64(const double4* __restrict__ tab,
                                                   const double4* __restrict__ rhs,
                                                   const GroupDesc* __restrict__ groups, int Npair,
                                                   int Nl, int nsets, int set,
                                                   double* __restrict__ a_equ,
                                                   double* __restrict__ a_pol) {
  const GroupDesc gd = groups[blockIdx.x];
  if (gd.nrows == 0) return;  // padding of the XCD schedule
  const int r = threadIdx.x;
  const bool valid = r < gd.nrows;
  const double4* t = tab + gd.off + r;
  // the set's four coefficient vectors of node pair np start at q[np * qs]
  const double4* q = rhs + ((int64_t)(gd.m * 2 + gd.P) * Npair * nsets + set) * 4;
  const int64_t qs = 4 * (int64_t)nsets;
  const int64_t stride = gd.nrows;
  double aer = 0, aei = 0, apr = 0, api = 0;
#pragma unroll 2
  for (int np = 0; np < Npair; ++np) {
    const double4 y = valid ? stream_load(t + np * stride) : make_double4(0, 0, 0, 0);
    const double4 qer = q[np * qs + 0], qei = q[np * qs + 1];
    const double4 qpr = q[np * qs + 2], qpi = q[np * qs + 3];
    aer = fma(y.x, qer.x, fma(y.y, qer.y, fma(y.z, qer.z, fma(y.w, qer.w, aer))));
    aei = fma(y.x, qei.x, fma(y.y, qei.y, fma(y.z, qei.z, fma(y.w, qei.w, aei))));
    apr = fma(y.x, qpr.x, fma(y.y, qpr.y, fma(y.z, qpr.z, fma(y.w, qpr.w, apr))));
    api = fma(y.x, qpi.x, fma(y.y, qpi.y, fma(y.z, qpi.z, fma(y.w, qpi.w, api))));
  }
  if (valid) {
    const int ell = gd.m + gd.P + 2 * (gd.i0 + r);
    const int64_t idx = hp_index(ell, gd.m, Nl) + (int64_t)set * Nl * (Nl + 1) / 2;
    a_equ[2 * idx] = aer;
    a_equ[2 * idx + 1] = aei;
    a_pol[2 * idx] = apr;
    a_pol[2 * idx + 1] = api;
  }
}

int launch_contract_set(Plan& pl, int nsets, int set, hipStream_t s) {
  Workspace& ws = pl.ws;
  if (pl.caller_tables) {
    k_contract64<<<pl.ngroups, 64, 0, s>>>(reinterpret_cast<const double4*>(pl.tab64),
                                           reinterpret_cast<const double4*>(ws.rhs), pl.groups,
                                           pl.Npair, pl.Nl, nsets, set, ws.a_equ, ws.a_pol);
    DSHT_LAUNCH_CHECK();
    return DSHT_OK;
  }
  const int64_t set_off = (int64_t)set * pl.Nlm;
  int nsplit = contract_splits(pl);
  dim3 cgrid(pl.ngroups, nsplit);
  if (pl.otf) {
    nsplit = std::min(kContractSplits, (2 * pl.Npair + kOtfNodes - 1) / kOtfNodes);
    k_contract_otf<<<dim3(pl.Nl, nsplit), kOtfWaves * 64, 0, s>>>(
        reinterpret_cast<const double2*>(pl.otf_rc), static_cast<const float*>(ws.rhs), pl.Nl,
        pl.Nx, pl.Npair, pl.dx, pl.Nlm, ws.a_der);
  } else if (pl.precision == DSHT_FP32) {
    k_contract_y<float><<<cgrid, 64, 0, s>>>(reinterpret_cast<const float4*>(pl.tab32y),
                                             static_cast<const float*>(ws.rhs), pl.groups, pl.Npair,
                                             pl.Nl, pl.Nlm, ws.a_der);
  } else {
    k_contract_y<double><<<cgrid, 64, 0, s>>>(reinterpret_cast<const double4*>(pl.tab64y),
                                              static_cast<const double*>(ws.rhs), pl.groups, pl.Npair,
                                              pl.Nl, pl.Nlm, ws.a_der);
  }
  DSHT_LAUNCH_CHECK();
  dim3 grid((pl.Nl + 255) / 256, pl.Nl);
  k_add_der<<<grid, 256, 0, s>>>(pl.Nl, pl.Nlm, nsplit, set_off, ws.a_der, ws.a_equ, ws.a_pol);
  DSHT_LAUNCH_CHECK();
  return DSHT_OK;
}

}  // namespace dsht
