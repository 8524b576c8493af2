// Operator SpMM and the fused block multi-shift CG (gfx950, wave64).
//
// Reference semantics: realworld/spmm.py:3-25 (csr_spmm), realworld/polar_gpu.py:63-80 (Op),
// realworld/polar_gpu.py:88-133 (msolve).  Design (not a translation of the CUDA kernel):
//  * Operator stored as diagonal + off-diagonal adjacency (value-free for unit weights).  A row is
//    processed by a group of L lanes, each lane owning CPL consecutive columns (16-byte loads);
//    the group loads L column indices in one coalesced load and broadcasts them with ds_bpermute.
//  * Rows longer than the chunk length are split into chunks processed by independent groups;
//    chunk partials are combined in chunk order by a finishing kernel (deterministic).
//  * Every reduction (d.v, |r|^2) is a fixed-order two-level sum: per-workgroup partials, then the
//    last workgroup to finish sums them in block order and runs the scalar recurrences of the
//    iteration (Jegerlehner zeta, alpha/beta of every pole, freezing, convergence) in FP64.
//  * Per iteration: opmul (+dot)  [-> longfin]  -> resid (+|r|^2, scalars)  -> update (d and the
//    shifted x/p of every live pole on the active rows).  A device flag makes launches after
//    convergence no-ops, so the host reads it only once per window of 8 iterations.
#include "common.hpp"

#include <mutex>
#include <vector>

namespace polar {

// ------------------------------------------------------------------------------------------------
// error state
static thread_local char g_err[1024] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* what, int line) {
  set_error((int)e, "%s failed at krylov/greedy line %d: %s", what, line, hipGetErrorString(e));
  return (int)e;
}

template <typename W>
__global__ __launch_bounds__(kBlock) void k_zero(W* __restrict__ p, int64_t words) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < words; i += stride) p[i] = W{};
}

int zero_async(void* p, size_t bytes, hipStream_t st) {
  if (bytes == 0) return 0;
  if (!p || bytes % 4 != 0) return set_error(POLAR_E_ARG, "zero_async: %zu bytes at %p", bytes, p);
  if (((uintptr_t)p & 15) == 0 && bytes % 16 == 0) {
    const int64_t words = (int64_t)(bytes / 16);
    const int grid = (int)std::min<int64_t>(65536, (words + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_zero<uint4>, dim3(grid), dim3(kBlock), 0, st, (uint4*)p, words);
  } else {
    const int64_t words = (int64_t)(bytes / 4);
    const int grid = (int)std::min<int64_t>(65536, (words + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_zero<uint32_t>, dim3(grid), dim3(kBlock), 0, st, (uint32_t*)p, words);
  }
  POLAR_LAUNCH_CHECK();
  return 0;
}

__device__ __forceinline__ int lower_bound_i32(const int32_t* __restrict__ a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------------------------------------------
// Order-independent reductions.  Every inner product and squared norm of the iteration is summed
// in double-double (error-free TwoSum with the error terms carried in a second double) and rounded
// once at the end, so its value is the correctly rounded sum of the (rounded) products whatever
// the grid shape or summation tree -- the CPU oracle computes the same numbers.
// Block reduction of per-lane column partials acc[CPL] (lanes with equal li own the same columns)
// into out[c] (hi) and out[R + c] (lo), c < R.
template <int CPL>
__device__ __forceinline__ void block_colsum(const Geo& g, const DD (&acc)[CPL], int R,
                                             double* __restrict__ out) {
  // lanes L, 2L, 4L, ... apart own the same columns: butterfly across the wave's groups, then
  // the kWaves wave results in LDS (fixed tree: the result does not depend on timing)
  __shared__ double red_hi[kWaves * 64 * 4];
  __shared__ double red_lo[kWaves * 64 * 4];
  DD v[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    v[k] = acc[k];
    for (int o = g.L; o < 64; o <<= 1) {
      DD w;
      w.hi = __shfl_xor(v[k].hi, o, 64);
      w.lo = __shfl_xor(v[k].lo, o, 64);
      v[k] = dd_merge(v[k], w);
    }
  }
  if (g.gi == 0) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      red_hi[(g.wave * 64 + g.li) * CPL + k] = v[k].hi;
      red_lo[(g.wave * 64 + g.li) * CPL + k] = v[k].lo;
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < R; c += kBlock) {
    const int li = c / CPL, k = c % CPL;
    DD s{0.0, 0.0};
    for (int w = 0; w < kWaves; ++w) {
      const int ix = (w * 64 + li) * CPL + k;
      s = dd_merge(s, DD{red_hi[ix], red_lo[ix]});
    }
    out[c] = s.hi;
    out[R + c] = s.lo;
  }
  __syncthreads();
}

// Column total of the double-double partial rows b in [b0, b1) u [b2, b3) (parts[b*2R + c] hi,
// parts[b*2R + R + c] lo) by one workgroup: strided double-double accumulation, a butterfly
// within each wave, then the waves in order; every thread gets the rounded total.
__device__ double col_total(const double* __restrict__ parts, int b0, int b1, int b2, int b3, int R, int c) {
  __shared__ double sh_hi[kWaves];
  __shared__ double sh_lo[kWaves];
  DD acc{0.0, 0.0};
  // partial rows [b0, b1) and [b2, b3)
  const int n01 = b1 - b0;
  const int nall = n01 + (b3 - b2);
#pragma unroll 4
  for (int k = threadIdx.x; k < nall; k += kBlock) {
    const int b = k < n01 ? b0 + k : b2 + (k - n01);
    const double* pb = parts + (int64_t)b * 2 * R;
    acc = dd_merge(acc, DD{pb[c], pb[R + c]});
  }
  acc = dd_group_sum(acc, 64);            // butterfly within the wave, then the waves in order
  if ((threadIdx.x & 63) == 0) {
    sh_hi[threadIdx.x >> 6] = acc.hi;
    sh_lo[threadIdx.x >> 6] = acc.lo;
  }
  __syncthreads();
  DD t{sh_hi[0], sh_lo[0]};
#pragma unroll
  for (int w = 1; w < kWaves; ++w) t = dd_merge(t, DD{sh_hi[w], sh_lo[w]});
  __syncthreads();
  return t.hi + t.lo;
}

// ------------------------------------------------------------------------------------------------
// Gather-accumulate acc[:] += sum_k w_k X[idx_k, c0:c0+CPL] over k in [beg, end), in k order.
// The group's L lanes load 8 consecutive column indices each (8L per index round); batches of 8
// gathers are issued before their ordered adds, so 8 row loads per lane are in flight.
// Every lane of the group takes part (the lanes share the index loads), whether its columns are in
// range (colok) or not.  TW: type of the stored weights (converted to T).
template <typename T, int CPL, bool WEIGHTED, typename TW = T>
__device__ __forceinline__ void gather_acc(const int32_t* __restrict__ idx, const TW* __restrict__ wt,
                                           const T* __restrict__ X, int64_t ldx, int beg, int end,
                                           const Geo& g, bool colok, int c0, T (&acc)[CPL]) {
  const int gbase = g.gi * g.L;
  for (int k0 = beg; k0 < end; k0 += 8 * g.L) {
    const int cnt = min(8 * g.L, end - k0);
    int myj[8];
    T myw[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = g.li * 8 + u;
      myj[u] = (t < cnt) ? idx[k0 + t] : 0;
      if (WEIGHTED) myw[u] = (t < cnt) ? (T)wt[k0 + t] : T(0);
    }
    for (int b = 0; 8 * b < cnt; ++b) {
      int j[8];
      T w[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        j[u] = __shfl(myj[u], gbase + b, 64);
        if (WEIGHTED) w[u] = __shfl(myw[u], gbase + b, 64);
      }
      if (colok) {
        const int nb = min(8, cnt - 8 * b);
        T x[8][CPL];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (u < nb) vload<T, CPL>(X + (int64_t)j[u] * ldx + c0, x[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (u < nb) {
#pragma unroll
            for (int e = 0; e < CPL; ++e) acc[e] += (WEIGHTED ? w[u] * x[u][e] : x[u][e]);
          }
      }
    }
  }
}

// Column groups: several systems on one graph, each a group of columns with its own poles, its
// own seed shift and its own added edges (patch entries tagged with the group).
struct ColGroups {
  const int* grp;        // [columns] column -> group, or NULL (one group: `shift` applies)
  const double* tau;     // [G][p]: tau[g*p] is group g's seed shift
  int p;
};

// y = (diag + shift) x - acc + patch(x) for a complete row; returns y in yv.  With column groups
// each column uses its group's shift and its group's patch entries (same arithmetic per column).
// finish_row_with: the row's own x (xv) and diagonal (dgd) already loaded.
template <typename T, int CPL>
__device__ __forceinline__ void finish_row_with(const polar_op& op, double shift, const T* __restrict__ X,
                                                int64_t ldx, int c0, int pidx, const T (&acc)[CPL],
                                                const T (&xv)[CPL], double dgd, T (&yv)[CPL], const ColGroups& cg) {
  int gcol[CPL];
  if (cg.grp) {
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      gcol[e] = cg.grp[c0 + e];
      const T dg = (T)(dgd + cg.tau[gcol[e] * cg.p]);
      yv[e] = (dg * xv[e]) - acc[e];
    }
  } else {
    const T dg = (T)(dgd + shift);
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      gcol[e] = 0;
      yv[e] = (dg * xv[e]) - acc[e];
    }
  }
  if (pidx >= 0) {
    const int pb = op.patch_ptr[pidx], pe = op.patch_ptr[pidx + 1];
    for (int q = pb; q < pe; ++q) {
      T xj[CPL];
      vload<T, CPL>(X + (int64_t)op.patch_cols[q] * ldx + c0, xj);
      const int pg = op.patch_grp ? op.patch_grp[q] : -1;
#pragma unroll
      for (int e = 0; e < CPL; ++e)
        if (pg < 0 || pg == gcol[e]) yv[e] = yv[e] + (xv[e] - xj[e]);
    }
  }
}

template <typename T, int CPL>
__device__ __forceinline__ void finish_row(const polar_op& op, double shift, const T* __restrict__ X,
                                           int64_t ldx, int row, int c0, int pidx, const T (&acc)[CPL],
                                           T (&xv)[CPL], T (&yv)[CPL], const ColGroups& cg) {
  vload<T, CPL>(X + (int64_t)row * ldx + c0, xv);
  const double dgd = (double)((const T*)op.diag)[row];
  finish_row_with<T, CPL>(op, shift, X, ldx, c0, pidx, acc, xv, dgd, yv, cg);
}

constexpr int kW = 8;            // window length of the blocked update (iterations between flushes)
constexpr int kBatch = kW;       // iterations launched between two convergence checks (one window)

// ------------------------------------------------------------------------------------------------
// CG state (device), laid out in the workspace.
struct CGState {
  // per column [R]
  double *rr, *stop, *a, *be, *ap, *bp;
  int *act, *dmask;
  int64_t* iters;
  // per pole x column [p][R]
  double *z, *zp, *zn, *aj, *bj, *h, *ip, *pn, *te, *tg, *shift;
  int *live, *m;
  int* flag;           // [2] any-active flags, indexed by iteration parity
  unsigned long long* stats;  // [3] sums over iterations of active columns and live (pole, column)
                              // pairs; (pole, column) pairs changed per window (blocked update)
  double* parts;       // [2*kMaxGrid][2][R] double-double workgroup partials
  int* colmap;         // [R] slot -> caller column (active-column compaction)
  int R, p;            // R: current (compacted) width
  int ld;              // row stride of vectors and P (the caller's width)
  int ident;           // colmap is the identity (no compaction yet)
  double tol;
  int freeze, track, trace;
  // Segments (batched form): S independent systems on disjoint row ranges of one block-diagonal
  // operator, each with its own poles and its own CG scalars for every column.  Per-column arrays
  // hold S*ld entries (segment g, column c at g*ld + c); per-pole arrays are [p][sld], sld = S*ld.
  // Blocks of every kernel are partitioned by segment (prefix tables over blockIdx), so partial
  // sums and finishers stay within a segment.  S == 1 needs no tables.
  int S, sld;
  int nb_a;                   // opmul blocks (the long-row finisher's partials follow them)
  const double* tau;          // [S][p] poles (tau[g*p] is segment g's seed shift)
  int* sflag;                 // [2][S] any-active flags per segment, by iteration parity
  const int* bA;              // [S+1] block prefix: k_opmul
  const int* bL;              // [S+1] k_longfin
  const int* bR;              // [S+1] row kernels with partials (k_init, k_resid)
  const int* bD;              // [S+1] k_update direction part
  const int* bS;              // [S+1] k_update shifted part
  const int64_t* seg_row;     // [S+1] row ranges
  const int64_t* seg_x;       // [S+1] long-row chunk unit ranges (entries of op.extra)
  const int64_t* seg_l;       // [S+1] long-row index ranges
  const int64_t* seg_a;       // [S+1] active-row ranges
  // Blocked shifted update (T > 1): the residuals of a window of T iterations are kept (ring of
  // T vectors) and X_j, P_j are updated once per window from per-(pole, column) coefficients:
  //   X_j <- X_j + cX[0] P_j + sum_t cX[t] r_t,   P_j <- cP[0] P_j + sum_t cP[t] r_t,
  // built by the finisher from the same aj, bj, zn the per-iteration update would apply.
  int T;
  double *cX, *cP;            // [p][T+1][sld]
  int* dirty;                 // [p][sld] (pole, column) changed in the current window
  int* pdirty;                // [p][S] pole j has a changed column of segment g
  // Residual replacement: the seed iterate x0 (n x ld, FP64) accumulated beside the recurrences;
  // every few windows the residual becomes w - (M + tau0) x0, computed in FP64 from the
  // operator values the solve uses.  NULL when off.  Per-iteration updates (T == 1): x0 += a d in
  // k_resid.  Blocked updates (T > 1): x0 advances once per window like a shifted block, from the
  // window coefficients of pole 0 (the seed pole: shift 0, zeta = 1, so its aj, bj are the seed's
  // alpha, beta) over the basis {d at the window start, r_1..r_T}.
  double* x0;
  // With the blocked update, x0 advances once per window: inside the window flush on the active rows
  // (whose residuals it already holds) and in k_seed_flush on the other rows; amask[i] != 0 marks
  // the active rows.
  const uint8_t* amask;
  // Column groups (one segment): G groups, cgrp[slot] = group of the column in that slot.
  int G;
  int* cgrp;
  // The operator kernels take the seed shift from tau[g * p] in device memory (segment g, 0 for one
  // system) instead of their `shift` argument: a solve never reads tau back to the host.
  int tau_dev;
  // POLAR_HISTORY (one segment): the finisher of iteration t keeps, per (pole, column slot), the
  // live flag and a_j, b_j, zeta_j of X_j += a_j P_j, P_j <- zeta_j r_{t+1} + b_j P_j at
  // [t][j][ld]; hist_T[0] = iterations that did work.  NULL otherwise.
  uint8_t* hm;
  double *haj, *hbj, *hzn;
  int* hist_T;
};

// Segment of block b of a kernel whose blocks are partitioned by the prefix table `pref`.
struct SegPart {
  int g, rank, nb;
};

__device__ __forceinline__ SegPart seg_part(const int* __restrict__ pref, int S, int b, int nb_total) {
  if (S <= 1) return SegPart{0, b, nb_total};
  int lo = 0, hi = S - 1;     // largest g with pref[g] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pref[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return SegPart{lo, b - pref[lo], pref[lo + 1] - pref[lo]};
}

__device__ __forceinline__ void seg_range(const int64_t* __restrict__ tab, int S, int g, int64_t whole,
                                          int64_t& lo, int64_t& hi) {
  if (S <= 1) {
    lo = 0;
    hi = whole;
  } else {
    lo = tab[g];
    hi = tab[g + 1];
  }
}

// The segment is still iterating (the whole-solve flag covers S == 1).
__device__ __forceinline__ bool seg_live(const CGState& s, int cur, int g) {
  return s.S <= 1 || s.sflag[cur * s.S + g] != 0;
}

// Partial-row ranges of segment g: the row kernels' blocks, and the operator's blocks followed by
// the long-row finisher's (for S == 1 the first nparts rows).
__device__ __forceinline__ void rows_parts(const CGState& s, int g, int nparts, int& b0, int& b1) {
  if (s.S <= 1) {
    b0 = 0;
    b1 = nparts;
  } else {
    b0 = s.bR[g];
    b1 = s.bR[g + 1];
  }
}

// Finisher after d.v (one workgroup per (segment, column)): a = act ? rr/dv : 0.
__global__ __launch_bounds__(kBlock) void k_alpha(CGState s, int nparts, int cur) {
  if (s.flag[cur] == 0) return;
  const int g = blockIdx.x / s.R, c = blockIdx.x - g * s.R;
  if (!seg_live(s, cur, g)) return;
  const int vc = g * s.ld + c;
  const bool act = s.act[vc] != 0;
  const double rr = s.rr[vc];
  double dv;
  if (s.S <= 1)
    dv = col_total(s.parts, 0, nparts, 0, 0, s.R, c);
  else
    dv = col_total(s.parts, s.bA[g], s.bA[g + 1], s.nb_a + s.bL[g], s.nb_a + s.bL[g + 1], s.R, c);
  if (threadIdx.x == 0) s.a[vc] = act ? rr / dv : 0.0;
}

// One (pole, column) pair's state for the scalar step, loaded before the column total so the
// loads overlap its reduction.
struct PoleIn {
  double z, zp, ip, pn, h, tg, te, sh;
  double cx[kW], cp[kW];           // window coefficients k < t (blocked update)
  int live;
};

__device__ __forceinline__ void pole_load(const CGState& s, int j, int idx, int vc, const double* shift, int t,
                                          PoleIn& q) {
  q.z = s.z[idx];
  q.zp = s.zp[idx];
  q.live = s.live[idx];
  q.sh = shift[j];
  if (s.track) {
    q.ip = s.ip[idx];
    q.pn = s.pn[idx];
    q.h = s.h[idx];
  }
  if (s.trace) {
    q.tg = s.tg[idx];
    q.te = s.te[idx];
  }
  if (s.T > 1) {
    const double* cx = s.cX + (int64_t)j * (s.T + 1) * s.sld + vc;
    const double* cp = s.cP + (int64_t)j * (s.T + 1) * s.sld + vc;
#pragma unroll
    for (int k = 0; k < kW; ++k)
      if (k < t) {
        q.cx[k] = cx[(int64_t)k * s.sld];
        q.cp[k] = cp[(int64_t)k * s.sld];
      }
  }
}

// Finisher after |rn|^2 (one workgroup per (segment, column), threads over poles): every scalar
// recurrence of polar_gpu.py:106-125 for the column, freezing, convergence and the any-active
// flags.  Everything that does not depend on this iteration's |r|^2 is loaded first.
__global__ __launch_bounds__(kBlock) void k_scalars(CGState s, int nparts, int cur, int wt, int hit) {
  if (s.flag[cur] == 0) return;
  const int g = blockIdx.x / s.R, c = blockIdx.x - g * s.R;
  if (!seg_live(s, cur, g)) return;
  const int vc = g * s.ld + c;
  const bool act = s.act[vc] != 0;
  const double a = s.a[vc], rr = s.rr[vc], bpv = s.bp[vc], apv = s.ap[vc], stop = s.stop[vc];
  const double* shift = s.shift + (s.G > 1 ? s.cgrp[c] : g) * s.p;
  const int t = wt + 1;              // window step of the blocked update
  PoleIn q0;
  const int j0 = threadIdx.x;
  if (j0 < s.p) pole_load(s, j0, j0 * s.sld + vc, vc, shift, t, q0);
  int b0, b1;
  rows_parts(s, g, nparts, b0, b1);
  const double rrn = col_total(s.parts, b0, b1, 0, 0, s.R, c);
  __shared__ int nlive_s;
  if (threadIdx.x == 0) nlive_s = 0;
  const double be = act ? rrn / rr : 0.0;
  const double cc = act ? a * bpv / apv : 0.0;
  const double nr = sqrt(rrn);
  __syncthreads();
  int nlive = 0;
  for (int j = j0; j < s.p; j += kBlock) {
    const int idx = j * s.sld + vc;
    PoleIn q;
    if (j == j0) q = q0;
    else pole_load(s, j, idx, vc, shift, t, q);
    const double z = q.z, zp = q.zp;
    const double zn = z / (1.0 + a * q.sh + cc * (1.0 - z / zp));
    const double fac = zn / z;
    const bool m = (q.live != 0) && act;
    const double aj = m ? a * fac : 0.0;
    const double bj = be * fac * fac;
    s.zn[idx] = zn;
    s.aj[idx] = aj;
    s.bj[idx] = bj;
    s.m[idx] = m ? 1 : 0;
    nlive += m ? 1 : 0;
    if (s.track) {
      s.h[idx] = q.h + (2.0 * aj * q.ip + aj * aj * q.pn);
      if (m) {
        s.ip[idx] = bj * (q.ip + aj * q.pn);
        s.pn[idx] = zn * zn * rrn + bj * bj * q.pn;
      }
    }
    if (s.trace) {
      s.te[idx] = q.te + aj * q.tg;
      if (m) s.tg[idx] = bj * q.tg;
    }
    if (m) {
      s.zp[idx] = z;
      s.z[idx] = zn;
    }
    if (s.T > 1 && m) {
      // window step t: X += aj P, P = zn r_t + bj P in the window basis {P_j, r_1..r_T}
      double* cx = s.cX + (int64_t)j * (s.T + 1) * s.sld + vc;
      double* cp = s.cP + (int64_t)j * (s.T + 1) * s.sld + vc;
#pragma unroll
      for (int k = 0; k < kW; ++k)
        if (k < t) {
          cx[(int64_t)k * s.sld] = q.cx[k] + aj * q.cp[k];
          cp[(int64_t)k * s.sld] = bj * q.cp[k];
        }
      cp[(int64_t)t * s.sld] = zn;
      s.dirty[idx] = 1;
      s.pdirty[j * s.S + g] = 1;
    }
    if (s.hm && hit >= 0) {
      const int64_t hx = ((int64_t)hit * s.p + j) * s.ld + c;
      s.hm[hx] = m ? 1 : 0;
      s.haj[hx] = aj;
      s.hbj[hx] = bj;
      s.hzn[hx] = zn;
    }
    if (s.freeze && m && fabs(zn) * nr <= stop) s.live[idx] = 0;
  }
  if (s.hist_T && hit >= 0 && blockIdx.x == 0 && threadIdx.x == 0) s.hist_T[0] = hit + 1;
  if (nlive) atomicAdd(&nlive_s, nlive);
  __syncthreads();
  if (threadIdx.x == 0) {
    s.dmask[vc] = act ? 1 : 0;
    s.be[vc] = be;
    bool still = false;
    if (act) {
      s.iters[vc] += 1;
      const bool done = nr <= stop;
      s.rr[vc] = rrn;
      s.ap[vc] = a;
      s.bp[vc] = be;
      if (done) s.act[vc] = 0; else still = true;
      atomicAdd(&s.stats[0], 1ull);
    }
    if (nlive_s) atomicAdd(&s.stats[1], (unsigned long long)nlive_s);
    if (still) {
      atomicOr(&s.flag[cur ^ 1], 1);
      if (s.S > 1) atomicOr(&s.sflag[(cur ^ 1) * s.S + g], 1);
    }
  }
}

// Finisher of the init pass: rr, stop, act and every per-pole scalar of (segment, column).
__global__ __launch_bounds__(kBlock) void k_init_fin(CGState s, int nparts) {
  const int g = blockIdx.x / s.R, c = blockIdx.x - g * s.R;
  const int vc = g * s.ld + c;
  int b0, b1;
  rows_parts(s, g, nparts, b0, b1);
  const double rr = col_total(s.parts, b0, b1, 0, 0, s.R, c);
  for (int j = threadIdx.x; j < s.p; j += kBlock) {
    const int idx = j * s.sld + vc;
    s.z[idx] = 1.0;
    s.zp[idx] = 1.0;
    s.zn[idx] = 1.0;
    s.aj[idx] = 0.0;
    s.bj[idx] = 0.0;
    s.live[idx] = rr > 0.0 ? 1 : 0;
    s.m[idx] = 0;
    s.h[idx] = 0.0;
    s.ip[idx] = 0.0;
    s.pn[idx] = rr;
    s.te[idx] = 0.0;
    s.tg[idx] = rr;
  }
  if (threadIdx.x == 0) {
    s.rr[vc] = rr;
    s.stop[vc] = s.tol * sqrt(rr);
    s.act[vc] = rr > 0.0 ? 1 : 0;
    s.dmask[vc] = 0;
    s.a[vc] = 0.0;
    s.be[vc] = 0.0;
    s.ap[vc] = 1.0;
    s.bp[vc] = 0.0;
    s.iters[vc] = 0;
    if (rr > 0.0) {
      atomicOr(&s.flag[0], 1);
      atomicOr(&s.flag[1], 1);
      if (s.S > 1) {
        atomicOr(&s.sflag[g], 1);
        atomicOr(&s.sflag[s.S + g], 1);
      }
    }
  }
}

// ------------------------------------------------------------------------------------------------
// A run of nr consecutive rows [ra, ra + nr) of the operator (nr < L), processed by one row group:
// lane li of the group holds ext = indptr[ra + li] (li <= nr).  The run's short rows are walked as
// one entry sequence, so a batch of 8 gathers spans row boundaries (a row of 4 entries shares its
// batch with the next row) and one index round serves several rows; each row still sums its
// entries in adjacency order, so the row values are those of the one-row-per-unit walk.  Long rows
// of the run are skipped (their chunks are units of the chunk list).  Every shuffle reads lanes of
// the same group under conditions that are uniform within the group.
template <typename T, int CPL, bool WEIGHTED, bool DOT, typename TW>
__device__ __forceinline__ void run_rows_walk(const polar_op& op, double shift, const T* __restrict__ X,
                                              int64_t ldx, T* __restrict__ Y, int64_t ldy, const Geo& g,
                                              bool colok, int c0, int ra, int nr, int ext, DD (&dacc)[CPL],
                                              int& pc, const ColGroups& cg) {
  const int gbase = g.gi * g.L;
  const int CH = op.chunk;
  // long rows of the run: bit i of lmask (i < nr)
  const int ext_next = __shfl(ext, gbase + min(g.li + 1, g.L - 1), 64);
  const bool isl = g.li < nr && (ext_next - ext) > CH;
  const uint64_t bal = __ballot(isl);
  const uint64_t lmask = g.L == 64 ? bal : ((bal >> gbase) & ((1ull << g.L) - 1));
  auto finish = [&](int rowabs, const T (&acc)[CPL], const T (&xv)[CPL], double dgd) {
    int pidx = -1;
    if (op.n_patch > 0) {
      if (pc < 0) pc = lower_bound_i32(op.patch_rows, op.n_patch, rowabs);
      while (pc < op.n_patch && op.patch_rows[pc] < rowabs) ++pc;
      if (pc < op.n_patch && op.patch_rows[pc] == rowabs) pidx = pc;
    }
    if (colok) {
      T yv[CPL];
      finish_row_with<T, CPL>(op, shift, X, ldx, c0, pidx, acc, xv, dgd, yv, cg);
      vstore<T, CPL>(Y + (int64_t)rowabs * ldy + c0, yv);
      if (DOT) {
#pragma unroll
        for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)xv[e] * (double)yv[e]);
      }
    }
  };
  int cr = 0;
  while (cr < nr) {
    const uint64_t rem = lmask >> cr;
    if (rem & 1ull) {               // skip the long rows at cr
      cr += (int)__builtin_ctzll(~rem);
      continue;
    }
    const int ce = rem == 0 ? nr : min(nr, cr + (int)__builtin_ctzll(rem));
    // rows [cr, ce) are short: entries [ext(cr), ext(ce)) in one sequence
    const int E0 = __shfl(ext, gbase + cr, 64);
    const int E1 = __shfl(ext, gbase + ce, 64);
    int row = cr;
    int rend = __shfl(ext, gbase + cr + 1, 64);
    T acc[CPL], xv[CPL];
    double dgd = 0.0;
#pragma unroll
    for (int e = 0; e < CPL; ++e) acc[e] = T(0);
    if (colok) {
      vload<T, CPL>(X + (int64_t)(ra + row) * ldx + c0, xv);
      dgd = (double)((const TW*)op.diag)[ra + row];
    }
    for (int k0 = E0; k0 < E1; k0 += 8 * g.L) {
      const int cnt = min(8 * g.L, E1 - k0);
      int myj[8];
      T myw[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = g.li * 8 + u;
        myj[u] = (t < cnt) ? op.indices[k0 + t] : 0;
        if (WEIGHTED) myw[u] = (t < cnt) ? (T)((const TW*)op.weights)[k0 + t] : T(0);
      }
      for (int b = 0; 8 * b < cnt; ++b) {
        int j[8];
        T w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          j[u] = __shfl(myj[u], gbase + b, 64);
          if (WEIGHTED) w[u] = __shfl(myw[u], gbase + b, 64);
        }
        const int nb = min(8, cnt - 8 * b);
        T x[8][CPL];
        if (colok) {
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (u < nb) vload<T, CPL>(X + (int64_t)j[u] * ldx + c0, x[u]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (u < nb) {
            const int k = k0 + 8 * b + u;
            while (k >= rend) {     // rows ending before entry k (empty rows included)
              finish(ra + row, acc, xv, dgd);
              ++row;
              rend = __shfl(ext, gbase + row + 1, 64);
#pragma unroll
              for (int e = 0; e < CPL; ++e) acc[e] = T(0);
              if (colok) {
                vload<T, CPL>(X + (int64_t)(ra + row) * ldx + c0, xv);
                dgd = (double)((const TW*)op.diag)[ra + row];
              }
            }
            if (colok) {
#pragma unroll
              for (int e = 0; e < CPL; ++e) acc[e] += (WEIGHTED ? w[u] * x[u][e] : x[u][e]);
            }
          }
        }
      }
    }
    // the last row with entries and any empty rows after it
    for (;;) {
      finish(ra + row, acc, xv, dgd);
      if (++row >= ce) break;
#pragma unroll
      for (int e = 0; e < CPL; ++e) acc[e] = T(0);
      if (colok) {
        vload<T, CPL>(X + (int64_t)(ra + row) * ldx + c0, xv);
        dgd = (double)((const TW*)op.diag)[ra + row];
      }
    }
    cr = ce;
  }
}

// Kernel A: Y = Op(X) for complete rows; long-row chunk partials to lpart; optional d.v partials.
// T: vectors; TW: the operator's stored values (diagonal, weights) -- T except for the residual
// replacement product, which applies the FP32 operator to the FP64 seed iterate in FP64.
// run > 1: the row units are runs of `run` consecutive rows (run < L; see run_rows_walk).
template <typename T, int CPL, bool WEIGHTED, bool DOT, typename TW = T, bool RUNS = false>
__global__ __launch_bounds__(kBlock) void k_opmul(polar_op op, double shift, const T* __restrict__ X,
                                                  int64_t ldx, T* __restrict__ Y, int64_t ldy, int R,
                                                  int L, T* __restrict__ lpart, CGState s, int cur, int run) {
  if (DOT && s.flag[cur] == 0) return;
  if constexpr (RUNS) {
    const SegPart sp = seg_part(s.bA, s.S, blockIdx.x, gridDim.x);
    if (DOT && !seg_live(s, cur, sp.g)) return;
    if (s.S > 1 || s.tau_dev) shift = s.tau[sp.g * s.p];
    int64_t r0, r1, x0, x1;
    seg_range(s.seg_row, s.S, sp.g, op.n, r0, r1);
    seg_range(s.seg_x, s.S, sp.g, op.n_extra, x0, x1);
    Geo g(L);
    const int c0 = g.li * CPL;
    const bool colok = c0 < R;
    const int64_t nx = x1 - x0;
    const int64_t nunits = nx + (r1 - r0 + run - 1) / run;
    const int gpb = kWaves * g.G;
    int64_t u = (int64_t)sp.rank * gpb + g.wave * g.G + g.gi;
    const int64_t stride = (int64_t)sp.nb * gpb;
    const int CH = op.chunk;
    const ColGroups cg{s.G > 1 ? s.cgrp : nullptr, s.tau, s.p};
    DD dacc[CPL];
#pragma unroll
    for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
    int pc = -1;
    // the next run's row extents, loaded one unit ahead
    int next_ext = 0;
    auto load_ext = [&](int64_t un) {
      if (un >= nx && un < nunits) {
        const int64_t ra = r0 + (un - nx) * run;
        const int nr = (int)min((int64_t)run, r1 - ra);
        if (g.li <= nr) next_ext = op.indptr[ra + g.li];
      }
    };
    load_ext(u);
    for (; u < nunits; u += stride) {
      const int64_t un = u + stride;
      if (u >= nx) {
        const int ext = next_ext;
        load_ext(un);
        const int ra = (int)(r0 + (u - nx) * run);
        const int nr = (int)min((int64_t)run, r1 - ra);
        run_rows_walk<T, CPL, WEIGHTED, DOT, TW>(op, shift, X, ldx, Y, ldy, g, colok, c0, ra, nr, ext, dacc, pc, cg);
      } else {
        load_ext(un);
        const int2 ex = reinterpret_cast<const int2*>(op.extra)[x0 + u];
        const int lidx = ex.x, chunk = ex.y;
        const int row = op.long_rows[lidx];
        const int rb = op.indptr[row], re = op.indptr[row + 1];
        const int beg = rb + chunk * CH;
        const int end = min(re, beg + CH);
        T acc[CPL];
#pragma unroll
        for (int e = 0; e < CPL; ++e) acc[e] = T(0);
        gather_acc<T, CPL, WEIGHTED, TW>(op.indices, (const TW*)op.weights, X, ldx, beg, end, g, colok, c0, acc);
        if (colok) vstore<T, CPL>(lpart + (int64_t)(op.long_base[lidx] + chunk) * R + c0, acc);
      }
    }
    if (DOT) block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)blockIdx.x * 2 * R);
    return;
  }
  const SegPart sp = seg_part(s.bA, s.S, blockIdx.x, gridDim.x);
  if (DOT && !seg_live(s, cur, sp.g)) return;
  if (s.S > 1 || s.tau_dev) shift = s.tau[sp.g * s.p];
  int64_t r0, r1, x0, x1;
  seg_range(s.seg_row, s.S, sp.g, op.n, r0, r1);
  seg_range(s.seg_x, s.S, sp.g, op.n_extra, x0, x1);
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  const int64_t nx = x1 - x0;
  const int64_t nunits = nx + (r1 - r0);
  const int gpb = kWaves * g.G;
  int64_t u = (int64_t)sp.rank * gpb + g.wave * g.G + g.gi;
  const int64_t stride = (int64_t)sp.nb * gpb;
  const int CH = op.chunk;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  int pc = -1;
  // row extents of the next row unit, loaded one unit ahead so that their latency overlaps the
  // current unit's gathers
  int nbeg = 0, nend = 0;
  if (u >= nx && u < nunits) {
    nbeg = op.indptr[r0 + (u - nx)];
    nend = op.indptr[r0 + (u - nx) + 1];
  }
  // the segment's units [0, nx): every chunk of its long rows (scheduled first: their latency
  // chains overlap the rest); units [nx, nx + rows): the remaining rows
  for (; u < nunits; u += stride) {
    int row, beg, end, lidx = -1, chunk = 0;
    const int64_t un = u + stride;
    if (u >= nx) {
      row = (int)(r0 + (u - nx));
      beg = nbeg;
      end = nend;
      if (un < nunits) {
        nbeg = op.indptr[r0 + (un - nx)];
        nend = op.indptr[r0 + (un - nx) + 1];
      }
      if (end - beg > CH) continue;    // long row: all its chunks are units of the chunk list
    } else {
      if (un >= nx && un < nunits) {
        nbeg = op.indptr[r0 + (un - nx)];
        nend = op.indptr[r0 + (un - nx) + 1];
      }
      const int2 ex = reinterpret_cast<const int2*>(op.extra)[x0 + u];
      lidx = ex.x;
      chunk = ex.y;
      row = op.long_rows[lidx];
      const int rb = op.indptr[row], re = op.indptr[row + 1];
      beg = rb + chunk * CH;
      end = min(re, beg + CH);
    }
    T acc[CPL];
#pragma unroll
    for (int e = 0; e < CPL; ++e) acc[e] = T(0);
    // a complete row's own x and diagonal are loaded before its gathers (their latency overlaps)
    T xv[CPL];
    double dgd = 0.0;
    if (lidx < 0 && colok) {
      vload<T, CPL>(X + (int64_t)row * ldx + c0, xv);
      dgd = (double)((const TW*)op.diag)[row];
    }
    gather_acc<T, CPL, WEIGHTED, TW>(op.indices, (const TW*)op.weights, X, ldx, beg, end, g, colok, c0, acc);
    if (lidx < 0) {
      int pidx = -1;
      if (op.n_patch > 0) {
        if (pc < 0) pc = lower_bound_i32(op.patch_rows, op.n_patch, row);
        while (pc < op.n_patch && op.patch_rows[pc] < row) ++pc;
        if (pc < op.n_patch && op.patch_rows[pc] == row) pidx = pc;
      }
      if (colok) {
        T yv[CPL];
        finish_row_with<T, CPL>(op, shift, X, ldx, c0, pidx, acc, xv, dgd, yv, ColGroups{s.G > 1 ? s.cgrp : nullptr, s.tau, s.p});
        vstore<T, CPL>(Y + (int64_t)row * ldy + c0, yv);
        if (DOT) {
#pragma unroll
          for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)xv[e] * (double)yv[e]);
        }
      }
    } else if (colok) {
      vstore<T, CPL>(lpart + (int64_t)(op.long_base[lidx] + chunk) * R + c0, acc);
    }
  }
  if (DOT) {
    block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)blockIdx.x * 2 * R);
  }
}

// Kernel L: combine long-row chunk partials in chunk order, finish the rows, d.v partials.
template <typename T, int CPL, bool DOT, typename TW = T>
__global__ __launch_bounds__(kBlock) void k_longfin(polar_op op, double shift, const T* __restrict__ X,
                                                    int64_t ldx, T* __restrict__ Y, int64_t ldy, int R,
                                                    int L, const T* __restrict__ lpart, CGState s,
                                                    int cur, int parts_off) {
  if (DOT && s.flag[cur] == 0) return;
  const SegPart sp = seg_part(s.bL, s.S, blockIdx.x, gridDim.x);
  if (DOT && !seg_live(s, cur, sp.g)) return;
  if (s.S > 1 || s.tau_dev) shift = s.tau[sp.g * s.p];
  int64_t l0, l1;
  seg_range(s.seg_l, s.S, sp.g, op.n_long, l0, l1);
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  const int gpb = kWaves * g.G;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  for (int64_t l = l0 + (int64_t)sp.rank * gpb + g.wave * g.G + g.gi; l < l1; l += (int64_t)sp.nb * gpb) {
    const int row = op.long_rows[l];
    const int deg = op.indptr[row + 1] - op.indptr[row];
    const int nch = (deg + op.chunk - 1) / op.chunk;
    int pidx = -1;
    if (op.n_patch > 0) {
      int pc = lower_bound_i32(op.patch_rows, op.n_patch, row);
      if (pc < op.n_patch && op.patch_rows[pc] == row) pidx = pc;
    }
    if (colok) {
      // the row's own x and diagonal are loaded before its chunk sums (their latency overlaps)
      T xv[CPL];
      vload<T, CPL>(X + (int64_t)row * ldx + c0, xv);
      const double dgd = (double)((const TW*)op.diag)[row];
      T acc[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) acc[e] = T(0);
      const T* base = lpart + (int64_t)op.long_base[l] * R + c0;
      for (int ch0 = 0; ch0 < nch; ch0 += 8) {
        const int nb = min(8, nch - ch0);
        T pv[8][CPL];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (u < nb) vload<T, CPL>(base + (int64_t)(ch0 + u) * R, pv[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (u < nb) {
#pragma unroll
            for (int e = 0; e < CPL; ++e) acc[e] += pv[u][e];
          }
      }
      T yv[CPL];
      finish_row_with<T, CPL>(op, shift, X, ldx, c0, pidx, acc, xv, dgd, yv, ColGroups{s.G > 1 ? s.cgrp : nullptr, s.tau, s.p});
      vstore<T, CPL>(Y + (int64_t)row * ldy + c0, yv);
      if (DOT) {
#pragma unroll
        for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)xv[e] * (double)yv[e]);
      }
    }
  }
  if (DOT) {
    block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)(parts_off + blockIdx.x) * 2 * R);
  }
}

// Kernel B: rn = r - a v (active columns) and |rn|^2 partials (finished by k_scalars).
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_resid(int64_t n, int R, int L, const T* rin, T* rout,
                                                  const T* __restrict__ v, const T* __restrict__ d, CGState s,
                                                  int cur) {
  // the finisher of this iteration raises flag[cur^1] again if a column is still active; when
  // converged, both parities stay 0 so every later launch is a no-op
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) s.flag[cur ^ 1] = 0;
    if (s.S > 1)
      for (int t = threadIdx.x; t < s.S; t += kBlock) s.sflag[(cur ^ 1) * s.S + t] = 0;
  }
  if (s.flag[cur] == 0) return;
  const SegPart sp = seg_part(s.bR, s.S, blockIdx.x, gridDim.x);
  if (!seg_live(s, cur, sp.g)) return;
  int64_t r0, r1;
  seg_range(s.seg_row, s.S, sp.g, n, r0, r1);
  const int vb = sp.g * s.ld;
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  if (colok) {
    T a[CPL];
    bool act[CPL];
    bool any = false;
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      act[e] = s.act[vb + c0 + e] != 0;
      a[e] = (T)s.a[vb + c0 + e];
      any |= act[e];
    }
    const int gpb = kWaves * g.G;
    for (int64_t i = r0 + (int64_t)sp.rank * gpb + g.wave * g.G + g.gi; i < r1; i += (int64_t)sp.nb * gpb) {
      T rv[CPL], vv[CPL];
      vload<T, CPL>(rin + i * s.ld + c0, rv);
      if (any) {
        vload<T, CPL>(v + i * s.ld + c0, vv);
#pragma unroll
        for (int e = 0; e < CPL; ++e)
          if (act[e]) rv[e] = rv[e] - (a[e] * vv[e]);
        if (s.x0 && s.T <= 1) {              // x0 += alpha d (FP64), for residual replacement
          T dv[CPL];
          vload<T, CPL>(d + i * s.ld + c0, dv);
          double* xp = s.x0 + i * s.ld + c0;
#pragma unroll
          for (int e = 0; e < CPL; ++e)
            if (act[e]) xp[e] = xp[e] + s.a[vb + c0 + e] * (double)dv[e];
        }
      }
      if (any || rin != rout) vstore<T, CPL>(rout + i * s.ld + c0, rv);
#pragma unroll
      for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)rv[e] * (double)rv[e]);
    }
  }
  block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)blockIdx.x * 2 * R);
}

// Kernel C: d = rn + be d (columns active at the start of the iteration) on all rows, and the
// shifted updates X_j += aj P_j; P_j = zn rn[active] + bj P_j (live poles) on the active rows.
template <typename T, int CPL, bool STAGED>
__global__ __launch_bounds__(kBlock) void k_update(int64_t n, int64_t nA, int p, int R, int L,
                                                   const T* __restrict__ r, T* __restrict__ d,
                                                   const int32_t* __restrict__ active,
                                                   T* __restrict__ P, T* __restrict__ Xo, int64_t sj,
                                                   int64_t si, int nb_d, int ps, CGState s, int cur) {
  if (s.flag[cur] == 0) return;
  Geo g(L);
  const int c0 = g.li * CPL;
  const int gpb = kWaves * g.G;
  if ((int)blockIdx.x < nb_d) {
    const SegPart sp = seg_part(s.bD, s.S, blockIdx.x, nb_d);
    if (!seg_live(s, cur, sp.g)) return;
    if (c0 >= R) return;     // no barrier in this part
    int64_t r0, r1;
    seg_range(s.seg_row, s.S, sp.g, n, r0, r1);
    const int vb = sp.g * s.ld;
    bool msk[CPL];
    T be[CPL];
    bool any = false;
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      msk[e] = s.dmask[vb + c0 + e] != 0;
      be[e] = (T)s.be[vb + c0 + e];
      any |= msk[e];
    }
    if (!any) return;
    for (int64_t i = r0 + (int64_t)sp.rank * gpb + g.wave * g.G + g.gi; i < r1; i += (int64_t)sp.nb * gpb) {
      T rv[CPL], dv[CPL];
      vload<T, CPL>(r + i * s.ld + c0, rv);
      vload<T, CPL>(d + i * s.ld + c0, dv);
#pragma unroll
      for (int e = 0; e < CPL; ++e)
        if (msk[e]) dv[e] = rv[e] + (be[e] * dv[e]);
      vstore<T, CPL>(d + i * s.ld + c0, dv);
    }
    return;
  }
  // ps > 1 (one segment, few active rows): the shifted blocks are ps copies of the row partition,
  // copy q taking the poles j with j % ps == q; each (pole, row, column) update is unchanged.
  SegPart sp;
  int part = 0;
  if (ps > 1) {
    const int nbr = (int)(gridDim.x - nb_d) / ps, bb = (int)blockIdx.x - nb_d;
    part = bb / nbr;
    sp = SegPart{0, bb - part * nbr, nbr};
  } else {
    sp = seg_part(s.bS, s.S, blockIdx.x - nb_d, gridDim.x - nb_d);
  }
  if (!seg_live(s, cur, sp.g)) return;     // uniform over the block, before any barrier
  int64_t a0, a1;
  seg_range(s.seg_a, s.S, sp.g, nA, a0, a1);
  const int vb = sp.g * s.ld;
  // Stage this block's view of the per-pole scalars in LDS and list the poles that are live in
  // some column (frozen poles are skipped entirely); falls back to global loads when p*R is large.
  constexpr int kStage = STAGED ? 1024 : 1;
  __shared__ double s_aj[kStage], s_bj[kStage], s_zn[kStage];
  __shared__ int s_m[kStage];
  __shared__ int s_jl[kStage];
  __shared__ int s_nj;
  const bool staged = STAGED;
  if (STAGED) {
    if (threadIdx.x == 0) s_nj = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < p * R; q += kBlock) {
      const int j = q / R, c = q - j * R;
      const int src = j * s.sld + vb + c;
      s_aj[q] = s.aj[src];
      s_bj[q] = s.bj[src];
      s_zn[q] = s.zn[src];
      s_m[q] = s.m[src];
    }
    __syncthreads();
    for (int j = part + threadIdx.x * ps; j < p; j += kBlock * ps) {
      bool any = false;
      for (int c = 0; c < R; ++c) any |= s_m[j * R + c] != 0;
      if (any) s_jl[atomicAdd(&s_nj, 1)] = j;   // pole order is irrelevant: poles are independent
    }
    __syncthreads();
  }
  const int nj = staged ? s_nj : (p - part + ps - 1) / ps;
  for (int64_t i = a0 + (int64_t)sp.rank * gpb + g.wave * g.G + g.gi; i < a1; i += (int64_t)sp.nb * gpb) {
    if (c0 >= R || nj == 0) break;
    T rn[CPL];
    vload<T, CPL>(r + (int64_t)active[i] * s.ld + c0, rn);
    for (int jj = 0; jj < nj; ++jj) {
      const int j = staged ? s_jl[jj] : part + jj * ps;
      bool mj[CPL];
      bool anym = false;
      T aj[CPL], bj[CPL], zn[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        if (staged) {
          const int q = j * R + c0 + e;
          mj[e] = s_m[q] != 0;
          aj[e] = (T)s_aj[q];
          bj[e] = (T)s_bj[q];
          zn[e] = (T)s_zn[q];
        } else {
          const int q = j * s.sld + vb + c0 + e;
          mj[e] = s.m[q] != 0;
          aj[e] = (T)s.aj[q];
          bj[e] = (T)s.bj[q];
          zn[e] = (T)s.zn[q];
        }
        anym |= mj[e];
      }
      if (!anym) continue;
      T* pp = P + ((int64_t)j * nA + i) * s.ld + c0;
      T* xrow = Xo + (int64_t)j * sj + i * si;
      T pv[CPL], xv[CPL];
      vload_nt<T, CPL>(pp, pv);                   // P and X rows: touched once per iteration
      if (s.ident) {
        vload_nt<T, CPL>(xrow + c0, xv);
      } else {
#pragma unroll
        for (int e = 0; e < CPL; ++e) xv[e] = mj[e] ? xrow[s.colmap[c0 + e]] : T(0);
      }
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        if (mj[e]) {
          xv[e] = xv[e] + (aj[e] * pv[e]);
          pv[e] = (zn[e] * rn[e]) + (bj[e] * pv[e]);
        }
      }
      if (s.ident) {
        vstore_nt<T, CPL>(xrow + c0, xv);
      } else {
#pragma unroll
        for (int e = 0; e < CPL; ++e)
          if (mj[e]) xrow[s.colmap[c0 + e]] = xv[e];
      }
      vstore_nt<T, CPL>(pp, pv);
    }
  }
}

// Blocked shifted update, once per window of kW iterations (after the window's last iteration):
// X_j += cX0 P_j + sum_t cXt r_t;  P_j = cP0 P_j + sum_t cPt r_t  on the active rows, for the
// poles changed in the window.  Each thread keeps its row's window residuals in registers.  The
// changed poles come from one flag per pole (pdirty) and are taken kFlushChunk at a time: a chunk's
// coefficients are staged in LDS (two buffers), all its P and X rows are loaded together, and the
// next chunk's coefficients are fetched into registers while the current chunk computes, so a
// chunk costs one memory round trip and one barrier instead of one per pole.
constexpr int kMaxPolesFlush = 1024;
constexpr int kFlushLanes = 16;    // lanes per row of a flush tile (column slices of 128 bytes)
constexpr int kFlushWaves = 16;    // 1024 threads: a tile of 64 rows shares each staged chunk
constexpr int kFlushBlock = 64 * kFlushWaves;
constexpr int kFlushChunk = 2;     // poles per staged chunk (2 and 4 measured alike; 2 spills nothing)

// Columns per lane of the flush: half the solve's 16-byte vectors (8 bytes per lane, 16 lanes per
// 128-byte row piece), so that a lane's window residuals and a chunk's P and X rows fit 4 waves per
// SIMD, one 1024-thread block per CU (4 FP32 columns per lane needed 182 registers, 2 waves).
inline int flush_cpl(int cpl) { return cpl >= 2 ? cpl / 2 : 1; }

template <typename T> struct Ring {
  const T* s[kW];                // r_t is in s[t % kW]
};

template <typename T, int CPL, bool IDENT>
__global__ __launch_bounds__(kFlushBlock, 4) void k_flush(int64_t nA, int p, int R, int L, int cw, Ring<T> ring,
                                                       const int32_t* __restrict__ active, T* __restrict__ P,
                                                       T* __restrict__ Xo, int64_t sj, int64_t si, CGState s,
                                                       const T* __restrict__ dvec, T* __restrict__ dsave) {
  // column slice blockIdx.y: columns [cbase, cbase + Rs) (cw = L * CPL columns per slice)
  const int cbase = blockIdx.y * cw;
  const int Rs = min(cw, R - cbase);
  const SegPart sp = seg_part(s.bS, s.S, blockIdx.x, gridDim.x);
  int64_t a0, a1;
  seg_range(s.seg_a, s.S, sp.g, nA, a0, a1);
  const int vb = sp.g * s.ld;
  // a slice is at most kFlushLanes lanes of CPL columns
  constexpr int kMaxR = kFlushLanes * CPL;
  constexpr int kNQ = (kW + 1) * kMaxR;            // coefficients per pole and slice (row stride kMaxR)
  constexpr int kStage = (kFlushChunk * kNQ + kFlushBlock - 1) / kFlushBlock;   // per thread and chunk
  __shared__ int s_jl[kMaxPolesFlush];
  __shared__ int s_nj;
  __shared__ T s_cx[2][kFlushChunk][kNQ], s_cp[2][kFlushChunk][kNQ];   // [buffer][pole][t * kMaxR + c]
  __shared__ int s_dirty[2][kFlushChunk][kMaxR];
  // residual replacement: pole 0's window coefficients in FP64, for the seed iterate x0
  __shared__ double s_c0[kNQ];
  const bool seed = dsave != nullptr;
  if (threadIdx.x == 0) s_nj = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < p; j += kFlushBlock)
    if (s.pdirty[j * s.S + sp.g]) s_jl[atomicAdd(&s_nj, 1)] = j;   // poles are independent: any order
  if (seed)
    for (int q = threadIdx.x; q < kNQ; q += kFlushBlock) {
      const int t = q / kMaxR, c = q % kMaxR;
      if (c < Rs) s_c0[q] = s.cX[(int64_t)t * s.sld + vb + cbase + c];
    }
  __syncthreads();
  const int nj = s_nj;
  if (nj == 0) return;                           // uniform over the block
  const int nch = (nj + kFlushChunk - 1) / kFlushChunk;
  // chunk ch's coefficients (FP64 in memory) and column flags into registers ...
  double gx[kStage], gp[kStage];
  int gd = 0;
  auto fetch = [&](int ch) {
    const int j0 = ch * kFlushChunk;
    const int npc = min(kFlushChunk, nj - j0);
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int q = threadIdx.x + u * kFlushBlock;
      const int pc = q / kNQ, t = (q % kNQ) / kMaxR, c = q % kMaxR;
      if (pc < npc && c < Rs) {
        const int64_t src = ((int64_t)s_jl[j0 + pc] * (kW + 1) + t) * s.sld + vb + cbase + c;
        gx[u] = s.cX[src];
        gp[u] = s.cP[src];
      }
    }
    if (threadIdx.x < npc * Rs) {
      const int pc = threadIdx.x / Rs, c = threadIdx.x - pc * Rs;
      gd = s.dirty[(int64_t)s_jl[j0 + pc] * s.sld + vb + cbase + c];
    }
  };
  // ... and from registers into LDS buffer b
  auto put = [&](int ch, int b) {
    const int npc = min(kFlushChunk, nj - ch * kFlushChunk);
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int q = threadIdx.x + u * kFlushBlock;
      const int pc = q / kNQ, c = q % kMaxR;
      if (pc < npc && c < Rs) {
        s_cx[b][pc][q % kNQ] = (T)gx[u];
        s_cp[b][pc][q % kNQ] = (T)gp[u];
      }
    }
    if (threadIdx.x < npc * Rs) {
      const int pc = threadIdx.x / Rs, c = threadIdx.x - pc * Rs;
      s_dirty[b][pc][c] = gd;
    }
  };
  Geo g(L);
  const int cl = g.li * CPL;                     // column within the slice
  const int c0 = cbase + cl;
  const bool colok = cl < Rs;
  // compacted columns (!IDENT): slot c0 + e holds original column colmap[c0 + e] of X
  int xcol[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) xcol[e] = IDENT || !colok ? c0 + e : s.colmap[c0 + e];
  const int64_t tile = (int64_t)kFlushWaves * g.G;
  for (int64_t base = a0 + (int64_t)sp.rank * tile; base < a1; base += (int64_t)sp.nb * tile) {
    T rr[kW][CPL];
    const int64_t row = base + g.wave * g.G + g.gi;
    const bool ok = colok && row < a1;
    if (ok) {
      const int64_t ar = active[row];
#pragma unroll
      for (int t = 1; t <= kW; ++t) vload<T, CPL>(ring.s[t % kW] + ar * s.ld + c0, rr[t - 1]);
      if (seed) {
        // seed iterate on this active row: x0 += cX0 d(window start) + sum_t cXt r_t with pole 0's
        // coefficients (the seed pole), the residuals already in registers; then the saved
        // direction becomes d, the next window's start
        T b0[CPL], dv[CPL];
        double xv0[CPL];
        vload<T, CPL>(dsave + ar * s.ld + c0, b0);
        vload<T, CPL>(dvec + ar * s.ld + c0, dv);
        double* xp = s.x0 + ar * s.ld + c0;
        vload<double, CPL>(xp, xv0);
#pragma unroll
        for (int e = 0; e < CPL; ++e) {
          __builtin_amdgcn_sched_barrier(0);     // one column's FP64 terms live at a time
          double inc = s_c0[cl + e] * (double)b0[e];
#pragma unroll
          for (int t = 1; t <= kW; ++t) inc = inc + s_c0[t * kMaxR + cl + e] * (double)rr[t - 1][e];
          xv0[e] = xv0[e] + inc;
        }
        vstore<double, CPL>(xp, xv0);
        vstore<T, CPL>(dsave + ar * s.ld + c0, dv);
      }
    }
    __syncthreads();                             // the previous tile's last buffer is consumed
    fetch(0);
    put(0, 0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
      const int b = ch & 1;
      const int j0 = ch * kFlushChunk;
      const int npc = min(kFlushChunk, nj - j0);
      if (ch + 1 < nch) fetch(ch + 1);           // in flight while this chunk computes
      // every P and X row of the chunk is requested before any is used; columns unchanged in the
      // window have identity coefficients: nothing to read or write
      bool mine[kFlushChunk];
      T* prow[kFlushChunk];
      T* xrow[kFlushChunk];
#pragma unroll
      for (int pc = 0; pc < kFlushChunk; ++pc) {
        // the pole is the same for the whole block: its row bases stay in scalar registers
        const int j = __builtin_amdgcn_readfirstlane(s_jl[j0 + min(pc, npc - 1)]);
        prow[pc] = P + ((int64_t)j * nA + row) * s.ld + c0;
        xrow[pc] = Xo + (int64_t)j * sj + row * si;
        int any = 0;
#pragma unroll
        for (int e = 0; e < CPL; ++e) any |= s_dirty[b][pc][cl + e];
        mine[pc] = ok && pc < npc && any != 0;
      }
      T pv[kFlushChunk][CPL], xv[kFlushChunk][CPL];
#pragma unroll
      for (int pc = 0; pc < kFlushChunk; ++pc) {
        if (mine[pc]) {
          vload_nt<T, CPL>(prow[pc], pv[pc]);     // P and X rows: touched once per window
          if constexpr (IDENT) {
            vload_nt<T, CPL>(xrow[pc] + c0, xv[pc]);
          } else {
#pragma unroll
            for (int e = 0; e < CPL; ++e) xv[pc][e] = xrow[pc][xcol[e]];
          }
        }
      }
#pragma unroll
      for (int pc = 0; pc < kFlushChunk; ++pc) {
        // one pole's coefficients live at a time (in two halves), so that the chunk's P and X rows
        // and the next chunk's coefficients fit beside them at 4 waves per SIMD
        __builtin_amdgcn_sched_barrier(0);
        if (!mine[pc]) continue;
        const T* cx = s_cx[b][pc];
        const T* cp = s_cp[b][pc];
        T ax[CPL], ap[CPL];
#pragma unroll
        for (int e = 0; e < CPL; ++e) {
          ax[e] = cx[cl + e] * pv[pc][e];
          ap[e] = cp[cl + e] * pv[pc][e];
        }
#pragma unroll
        for (int t = 1; t <= kW; ++t) {
          if (t == kW / 2 + 1) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int e = 0; e < CPL; ++e) {
            ax[e] = ax[e] + cx[t * kMaxR + cl + e] * rr[t - 1][e];
            ap[e] = ap[e] + cp[t * kMaxR + cl + e] * rr[t - 1][e];
          }
        }
#pragma unroll
        for (int e = 0; e < CPL; ++e) {
          xv[pc][e] = xv[pc][e] + ax[e];
          pv[pc][e] = ap[e];
        }
        if constexpr (IDENT) {
          vstore_nt<T, CPL>(xrow[pc] + c0, xv[pc]);
        } else {
#pragma unroll
          for (int e = 0; e < CPL; ++e) xrow[pc][xcol[e]] = xv[pc][e];
        }
        vstore_nt<T, CPL>(prow[pc], pv[pc]);
      }
      if (ch + 1 < nch) put(ch + 1, b ^ 1);      // b ^ 1 was last read before the previous barrier
      __syncthreads();
    }
  }
}

// Start a window: identity coefficients (cP0 = 1, the rest 0), nothing dirty.  Counts the dirty
// (pole, column) pairs of the window that ends (stats[2]).
__global__ void k_coef_reset(CGState s) {
  const int64_t tot = (int64_t)s.p * s.sld;
  unsigned long long nd = 0;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx / s.sld, vc = idx - j * s.sld;
    double* cx = s.cX + j * (s.T + 1) * s.sld + vc;
    double* cp = s.cP + j * (s.T + 1) * s.sld + vc;
    for (int t = 0; t <= s.T; ++t) {
      cx[(int64_t)t * s.sld] = 0.0;
      cp[(int64_t)t * s.sld] = t == 0 ? 1.0 : 0.0;
    }
    nd += s.dirty[idx] != 0;
    s.dirty[idx] = 0;
  }
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)s.p * s.S;
       idx += (int64_t)gridDim.x * blockDim.x)
    s.pdirty[idx] = 0;
  if (nd) atomicAdd(&s.stats[2], nd);
}

// Residual replacement, second step: r = w - y on the rows of live segments for active columns, where
// y = (M + tau0) x0 was formed in FP64 by the operator kernel from the solve's own operator values;
// |r|^2 partials for k_replace_fin.  W is in the caller's column order: after active-column
// compaction slot c holds the caller's column colmap[c].
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_replace_apply(int64_t n, const T* __restrict__ W, int64_t ldw,
                                                          const double* __restrict__ y, T* __restrict__ r, int R,
                                                          int L, CGState s, int cur) {
  const SegPart sp = seg_part(s.bR, s.S, blockIdx.x, gridDim.x);
  if (s.flag[cur] == 0 || !seg_live(s, cur, sp.g)) return;
  int64_t r0, r1;
  seg_range(s.seg_row, s.S, sp.g, n, r0, r1);
  const int vb = sp.g * s.ld;
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  if (colok) {
    bool act[CPL];
    int wcol[CPL];
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      act[e] = s.act[vb + c0 + e] != 0;
      wcol[e] = s.ident ? c0 + e : s.colmap[c0 + e];
    }
    const int gpb = kWaves * g.G;
    for (int64_t i = r0 + (int64_t)sp.rank * gpb + g.wave * g.G + g.gi; i < r1; i += (int64_t)sp.nb * gpb) {
      T rv[CPL];
      double yv[CPL];
      vload<T, CPL>(r + i * s.ld + c0, rv);
      vload<double, CPL>(y + i * s.ld + c0, yv);
#pragma unroll
      for (int e = 0; e < CPL; ++e)
        if (act[e]) rv[e] = (T)((double)W[i * ldw + wcol[e]] - yv[e]);
      vstore<T, CPL>(r + i * s.ld + c0, rv);
#pragma unroll
      for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)rv[e] * (double)rv[e]);
    }
  }
  block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)blockIdx.x * 2 * R);
}

// Seed iterate of the blocked update on the rows outside the active set (the window flush advances
// it on the active rows), once per window before the coefficients are reset:
// x0 += cX0 dsave + sum_t cXt r_t with pole 0's window coefficients (pole 0 is the seed pole, so
// its coefficients describe x0 and d in the window basis {d at the window start, r_1..r_T}); then
// dsave = d, the next window's starting direction.  FP64 x0.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_seed_flush(int64_t n, int R, int L, Ring<T> ring, const T* __restrict__ d,
                                                       T* __restrict__ dsave, CGState s) {
  const SegPart sp = seg_part(s.bD, s.S, blockIdx.x, gridDim.x);
  int64_t r0, r1;
  seg_range(s.seg_row, s.S, sp.g, n, r0, r1);
  const int vb = sp.g * s.ld;
  Geo g(L);
  const int c0 = g.li * CPL;
  if (c0 >= R) return;     // no barrier in this kernel
  double cx[kW + 1][CPL];
  bool any = false;
#pragma unroll
  for (int t = 0; t <= kW; ++t)
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      cx[t][e] = s.cX[(int64_t)t * s.sld + vb + c0 + e];      // pole 0: rows [0, T] of cX
      any |= cx[t][e] != 0.0;
    }
  const int gpb = kWaves * g.G;
  for (int64_t i = r0 + (int64_t)sp.rank * gpb + g.wave * g.G + g.gi; i < r1; i += (int64_t)sp.nb * gpb) {
    if (s.amask[i]) continue;
    T dv[CPL];
    vload<T, CPL>(d + i * s.ld + c0, dv);
    if (any) {
      T b[CPL];
      double xv[CPL];
      double* xp = s.x0 + i * s.ld + c0;
      vload<double, CPL>(xp, xv);
      vload<T, CPL>(dsave + i * s.ld + c0, b);
      double inc[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) inc[e] = cx[0][e] * (double)b[e];
#pragma unroll
      for (int t = 1; t <= kW; ++t) {
        // steps the window did not take (convergence, or the end of a history solve) have zero
        // weight and a residual that may never have been written: skipped
        bool nz = false;
#pragma unroll
        for (int e = 0; e < CPL; ++e) nz |= cx[t][e] != 0.0;
        if (!nz) continue;
        vload<T, CPL>(ring.s[t % kW] + i * s.ld + c0, b);
#pragma unroll
        for (int e = 0; e < CPL; ++e)
          if (cx[t][e] != 0.0) inc[e] = inc[e] + cx[t][e] * (double)b[e];
      }
#pragma unroll
      for (int e = 0; e < CPL; ++e) xv[e] = xv[e] + inc[e];
      vstore<double, CPL>(xp, xv);
    }
    vstore<T, CPL>(dsave + i * s.ld + c0, dv);
  }
}

// After a replacement: rr of every active (segment, column) is the replaced residual's norm.
__global__ __launch_bounds__(kBlock) void k_replace_fin(CGState s, int nparts, int cur) {
  if (s.flag[cur] == 0) return;
  const int g = blockIdx.x / s.R, c = blockIdx.x - g * s.R;
  if (!seg_live(s, cur, g)) return;
  const int vc = g * s.ld + c;
  int b0, b1;
  rows_parts(s, g, nparts, b0, b1);
  const double rr = col_total(s.parts, b0, b1, 0, 0, s.R, c);
  if (threadIdx.x == 0 && s.act[vc]) s.rr[vc] = rr;
}

// Init: r = d = W[:, :R]; |w|^2 partials; last block initialises every scalar.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_init(int64_t n, int R, int L, const T* __restrict__ W,
                                                 int64_t ldw, T* __restrict__ r, T* __restrict__ d,
                                                 CGState s) {
  const SegPart sp = seg_part(s.bR, s.S, blockIdx.x, gridDim.x);
  int64_t r0, r1;
  seg_range(s.seg_row, s.S, sp.g, n, r0, r1);
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  if (colok) {
    const int gpb = kWaves * g.G;
    for (int64_t i = r0 + (int64_t)sp.rank * gpb + g.wave * g.G + g.gi; i < r1; i += (int64_t)sp.nb * gpb) {
      T wv[CPL];
      vload<T, CPL>(W + i * ldw + c0, wv);
      vstore<T, CPL>(r + i * s.ld + c0, wv);
      vstore<T, CPL>(d + i * s.ld + c0, wv);
#pragma unroll
      for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)wv[e] * (double)wv[e]);
    }
  }
  block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)blockIdx.x * 2 * R);
}

// P_j = W[active, :R] for every pole; X_j = 0.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_init_shift(int64_t nA, int p, int R, int L,
                                                       const T* __restrict__ W, int64_t ldw,
                                                       const int32_t* __restrict__ active,
                                                       T* __restrict__ P, T* __restrict__ Xo,
                                                       int64_t sj, int64_t si) {
  Geo g(L);
  const int c0 = g.li * CPL;
  if (c0 >= R) return;
  const int gpb = kWaves * g.G;
  for (int64_t i = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; i < nA; i += (int64_t)gridDim.x * gpb) {
    T wv[CPL], zero[CPL];
    vload<T, CPL>(W + (int64_t)active[i] * ldw + c0, wv);
#pragma unroll
    for (int e = 0; e < CPL; ++e) zero[e] = T(0);
    for (int j = 0; j < p; ++j) {             // streaming writes: the blocks are far larger than L2
      vstore_nt<T, CPL>(P + ((int64_t)j * nA + i) * R + c0, wv);
      vstore_nt<T, CPL>(Xo + (int64_t)j * sj + i * si + c0, zero);
    }
  }
}

// shift[g*p + j] = tau[g*p + j] - tau[g*p] (every segment's poles relative to its seed shift)
__global__ void k_shift(const double* __restrict__ tau, int p, int S, double* __restrict__ shift) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p * S) shift[i] = tau[i] - tau[(i / p) * p];
}

// ------------------------------------------------------------------------------------------------
// Host side
// Live kernel profiling (HIP events around each launch of the CG iteration), off by default.
// Classes 0..5: k_opmul, k_longfin, k_alpha, k_resid, k_scalars, k_update; class 6 counts solves
// and, in `bytes`, the sum over iterations of live (pole, column) pairs; class 7 counts every
// iteration that did work (in mode 3 classes 0..5 cover only the timed sample of them).
struct ProfClass {
  double ms = 0.0;
  int64_t launches = 0;
  double bytes = 0.0;
};
constexpr int kEv = 7;        // events per iteration (6 kernels)
constexpr int kProfClasses = kEv + 1;
static ProfClass g_prof[kProfClasses];
static int g_prof_on = 0;     // 0 off; 1 every class; 2 k_opmul, every iteration; 3 k_opmul, 1 in 8

struct Plan {
  int cpl, L, G;
  int run;                         // rows per operator work unit (1: one row per unit)
  int grid_a, grid_l, grid_rows;   // blocks of the kernels that write partial sums
  int nb_d, nb_s;                  // k_update: direction part, shifted part (or the window flush)
  int Lf;                          // lanes per row of the shifted part (flush: column slices)
  int fin;                         // finisher grid: one workgroup per (segment, column)
  std::vector<int> tab;            // S > 1: prefix tables bA, bL, bR, bD, bS (S+1 entries each)
};

// Segment ranges on the host (S == 1: the whole operator).
struct Segs {
  int S = 1;
  std::vector<int64_t> row, x, l, a;   // [S+1]: rows, long-row chunk units, long rows, active rows
  int G = 1;                           // column groups (S == 1 only)
  std::vector<int> col_group;          // [R] when G > 1
};

constexpr int kBatchGrid = 4096;   // default block budget of a batched kernel with partial sums
constexpr int kUpdGrid = 65536;    // block budget of the stream kernels without partial sums
constexpr int kSplitGrid = 512;    // k_update's shifted part splits its poles below this many blocks

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// Blocks of one kernel over the segments: units[g] work units of segment g (0: none or finished),
// gpb units per block and pass, about `cap` blocks in total (at least one per segment with work).
// pref[g]..pref[g+1] are segment g's blocks.
static void partition(const std::vector<int64_t>& units, int64_t gpb, int64_t cap, int* pref) {
  int64_t tot = 0;
  for (int64_t u : units) tot += u;
  const int64_t upg = std::max<int64_t>(1, (tot + gpb * cap - 1) / (gpb * cap));
  pref[0] = 0;
  for (size_t g = 0; g < units.size(); ++g) {
    const int64_t b = units[g] > 0 ? (units[g] + gpb * upg - 1) / (gpb * upg) : 0;
    pref[g + 1] = pref[g] + (int)b;
  }
}

static int batch_cap(const polar_op* op) { return op->grid_max > 0 ? op->grid_max : kBatchGrid; }

// Block budget of the single-system operator kernel.  On operators of a million rows and more the
// gathers gain from more resident waves (n=4M, R=128 FP32: 5.69 ms at 1,024 blocks, 5.37 ms at
// 16,384); the other partial-sum kernels keep kMaxGrid, so only the d.v total reads more partials.
constexpr int kOpGridLarge = 16384;
constexpr int64_t kLargeRows = 1 << 20;
static int64_t op_grid_cap(const polar_op* op) {
  if (op->grid_max > 0) return op->grid_max;
  return op->n + op->n_extra >= kLargeRows ? kOpGridLarge : kMaxGrid;
}

// Rows per operator work unit: op->run_rows (0 = one row per unit).  A run's row extents live in
// the lanes of its row group, so a run has fewer rows than the group has lanes.
static int run_for(const polar_op* op, int L) {
  return std::max(1, std::min(std::max(op->run_rows, 1), L - 1));
}

// Work units of rows [0, rows) at `run` rows per unit.
static int64_t row_units(int64_t rows, int run) { return (rows + run - 1) / run; }

static Plan make_plan(const polar_op* op, int64_t R, int cpl, const Segs* sg = nullptr,
                      const std::vector<char>* live = nullptr, int Tw = 1) {
  Plan pl;
  pl.cpl = cpl;
  pl.L = lanes_for(R, cpl);
  pl.G = 64 / pl.L;
  pl.run = run_for(op, pl.L);
  const int64_t gpb = (int64_t)kWaves * pl.G;
  // blocked update: the shifted part is the window flush, whose tiles are column slices of at
  // most kFlushLanes lanes per row (more rows per tile); otherwise k_update's full-width rows
  pl.Lf = Tw > 1 ? std::min(lanes_for(R, flush_cpl(cpl)), kFlushLanes) : pl.L;
  const int64_t gpb_s = (int64_t)(Tw > 1 ? kFlushWaves : kWaves) * (64 / pl.Lf);
  const int S = sg ? sg->S : 1;
  pl.fin = (int)(S * R);
  if (S == 1) {
    const int64_t units = row_units(op->n, pl.run) + op->n_extra;
    const int64_t nA = sg ? sg->a[1] : 0;
    const int64_t cap = op->grid_max > 0 ? op->grid_max : kMaxGrid;
    pl.grid_a = (int)std::min<int64_t>(op_grid_cap(op), std::max<int64_t>(1, (units + gpb - 1) / gpb));
    pl.grid_l = (int)std::min<int64_t>(cap, std::max<int64_t>(1, ((int64_t)op->n_long + gpb - 1) / gpb));
    pl.grid_rows = (int)std::min<int64_t>(cap, std::max<int64_t>(1, (op->n + gpb - 1) / gpb));
    pl.nb_d = (int)std::min<int64_t>(kUpdGrid, std::max<int64_t>(1, (op->n + gpb - 1) / gpb));
    pl.nb_s = (int)std::min<int64_t>(kUpdGrid, std::max<int64_t>(1, (nA + gpb_s - 1) / gpb_s));
    return pl;
  }
  const int64_t cap = batch_cap(op);
  std::vector<int64_t> ua(S), ul(S), ur(S), us(S);
  for (int g = 0; g < S; ++g) {
    const bool on = !live || (*live)[g];
    ur[g] = on ? sg->row[g + 1] - sg->row[g] : 0;
    ua[g] = on ? row_units(ur[g], pl.run) + (sg->x[g + 1] - sg->x[g]) : 0;
    ul[g] = on ? sg->l[g + 1] - sg->l[g] : 0;
    us[g] = on ? sg->a[g + 1] - sg->a[g] : 0;
  }
  pl.tab.assign(5 * (S + 1), 0);
  int* t = pl.tab.data();
  partition(ua, gpb, cap, t);
  partition(ul, gpb, cap, t + (S + 1));
  partition(ur, gpb, cap, t + 2 * (S + 1));
  partition(ur, gpb, kUpdGrid, t + 3 * (S + 1));
  partition(us, gpb_s, kUpdGrid, t + 4 * (S + 1));
  pl.grid_a = t[S];
  pl.grid_l = t[2 * (S + 1) - 1];
  pl.grid_rows = t[3 * (S + 1) - 1];
  pl.nb_d = t[4 * (S + 1) - 1];
  pl.nb_s = t[5 * (S + 1) - 1];
  return pl;
}

// Partial-sum rows the workspace holds for any plan of this operator and segmentation.
static int64_t parts_rows(const polar_op* op, int S) {
  if (S == 1) return std::max<int64_t>(op_grid_cap(op), kMaxGrid) + std::max(kMaxGrid, op->grid_max);
  return 2 * ((int64_t)batch_cap(op) + S);
}

static int pick_cpl(int dtype, int64_t R, std::initializer_list<std::pair<const void*, int64_t>> ptrs,
                    int cpl_max = 0) {
  const int esz = elem_size(dtype);
  int full = 16 / esz;
  if (cpl_max > 0) full = std::min(full, cpl_max);
  for (int cpl = full; cpl > 1; cpl /= 2) {
    bool ok = (R % cpl) == 0 && (R + cpl - 1) / cpl <= 64;
    for (auto& pr : ptrs) {
      const int64_t bytes_ld = pr.second * esz;
      ok = ok && (((uintptr_t)pr.first) % (cpl * esz) == 0) && (bytes_ld % (cpl * esz) == 0);
    }
    if (ok) return cpl;
  }
  return 1;
}

static int check_op(const polar_op* op) {
  if (!op || op->n <= 0 || !op->indptr || !op->diag || (op->dtype != POLAR_F64 && op->dtype != POLAR_F32))
    return set_error(POLAR_E_ARG, "invalid polar_op");
  if (op->weighted && !op->weights) return set_error(POLAR_E_ARG, "weighted op without weights");
  if (op->n_long > 0 && (!op->long_rows || !op->long_base)) return set_error(POLAR_E_ARG, "long-row schedule missing");
  if (op->n_extra > 0 && !op->extra) return set_error(POLAR_E_ARG, "extra chunk list missing");
  if (op->n_patch > 0 && (!op->patch_rows || !op->patch_ptr || !op->patch_cols))
    return set_error(POLAR_E_ARG, "patch arrays missing");
  if (op->n > INT32_MAX) return set_error(POLAR_E_ARG, "n exceeds int32 row ids");
  if (op->chunk < 8) return set_error(POLAR_E_ARG, "long-row chunk length must be >= 8");
  if (op->cpl_max < 0 || op->cpl_max > 4 || (op->cpl_max & (op->cpl_max - 1)))
    return set_error(POLAR_E_ARG, "cpl_max must be 0, 1, 2 or 4");
  if (op->grid_max < 0 || op->grid_max > 65536) return set_error(POLAR_E_ARG, "grid_max out of range");
  if (op->run_rows < 0 || op->run_rows > 63) return set_error(POLAR_E_ARG, "run_rows must be in [0, 63]");
  return 0;
}

// Workspace carve-up.
struct Carver {
  char* base;
  int64_t off = 0;
  explicit Carver(void* b) : base((char*)b) {}
  template <typename X> X* take(int64_t count) {
    off = align_up(off, 256);
    X* p = base ? (X*)(base + off) : nullptr;
    off += count * (int64_t)sizeof(X);
    return p;
  }
};

// Device state of one (batched) solve: S segments of R columns, p poles, nA active rows in total.
template <typename T>
struct SolveMem {
  CGState s{};
  int* perm;
  T *ring, *d, *v, *t2, *P, *lpart;   // ring: Tw residual vectors (Tw = 1: the residual itself)
  T* dsave;            // residual replacement with the blocked update: d at the window start
  double *y64, *lpart64;   // residual replacement: the FP64 product and its long-row partials
  int* dtab;           // 5 (S+1) block prefix tables
  int64_t* dseg;       // 4 (S+1) segment ranges
  // POLAR_HISTORY: ring holds r_0 .. r_tcap; `cur` is the residual that replaced a stored one (the
  // next iteration reads it); coef[c][j][t] is the weight of r_t in X_j (formed on demand)
  int64_t tcap;
  T* cur;
  double* coef;
};

template <typename T>
static int64_t carve_solve(void* base, const polar_op* op, int S, int64_t R, int p, int64_t nA, int Tw,
                           bool repl, SolveMem<T>& m, int64_t tcap = 0) {
  Carver cv(base);
  m.tcap = tcap;
  CGState& s = m.s;
  const int64_t C = (int64_t)S * R;
  s.rr = cv.take<double>(C); s.stop = cv.take<double>(C); s.a = cv.take<double>(C);
  s.be = cv.take<double>(C); s.ap = cv.take<double>(C); s.bp = cv.take<double>(C);
  // the two any-active flags and the per-column active mask are adjacent: the host reads them
  // with one copy at each convergence check
  int* fa = cv.take<int>(2 + C);
  s.flag = fa;
  s.act = fa + 2;
  s.dmask = cv.take<int>(C); s.iters = cv.take<int64_t>(C);
  const int64_t pr = (int64_t)p * C;
  s.z = cv.take<double>(pr); s.zp = cv.take<double>(pr); s.zn = cv.take<double>(pr);
  s.aj = cv.take<double>(pr); s.bj = cv.take<double>(pr); s.h = cv.take<double>(pr);
  s.ip = cv.take<double>(pr); s.pn = cv.take<double>(pr); s.te = cv.take<double>(pr);
  s.tg = cv.take<double>(pr); s.shift = cv.take<double>((int64_t)p * std::max<int64_t>(S, R));
  s.live = cv.take<int>(pr); s.m = cv.take<int>(pr);
  s.stats = cv.take<unsigned long long>(3);
  s.parts = cv.take<double>(parts_rows(op, S) * 2 * R);
  if (Tw > 1) {
    s.cX = cv.take<double>(pr * (Tw + 1));
    s.cP = cv.take<double>(pr * (Tw + 1));
    s.dirty = cv.take<int>(pr);
    s.pdirty = cv.take<int>((int64_t)p * S);
  }
  if (repl) s.x0 = cv.take<double>(op->n * R);
  // blocked update with replacement: the active-row mask, and the saved direction of the rows
  // outside the active set (when there are any)
  s.amask = (repl && Tw > 1) ? cv.take<uint8_t>(align_up(op->n, 16)) : nullptr;
  m.dsave = (repl && Tw > 1) ? cv.take<T>(op->n * R) : nullptr;
  // the replacement product (M + tau0) x0 in FP64, and its long-row chunk partials
  m.y64 = repl ? cv.take<double>(op->n * R) : nullptr;
  m.lpart64 = repl ? cv.take<double>((int64_t)op->n_partial * R) : nullptr;
  s.cgrp = cv.take<int>(R);
  s.colmap = cv.take<int>(R);
  s.sflag = cv.take<int>(2 * S);
  m.dtab = cv.take<int>(5 * (S + 1));
  m.dseg = cv.take<int64_t>(4 * (S + 1));
  m.perm = cv.take<int>(R);
  const int64_t n = op->n;
  if (tcap > 0) {
    s.hist_T = cv.take<int>(1);
    s.hm = cv.take<uint8_t>(tcap * p * R);
    s.haj = cv.take<double>(tcap * p * R);
    s.hbj = cv.take<double>(tcap * p * R);
    s.hzn = cv.take<double>(tcap * p * R);
    m.coef = cv.take<double>(R * p * (tcap + 1));
    m.cur = cv.take<T>(n * R);
  } else {
    s.hist_T = nullptr;
    s.hm = nullptr;
    s.haj = s.hbj = s.hzn = nullptr;
    m.coef = nullptr;
    m.cur = nullptr;
  }
  m.ring = cv.take<T>(n * R * (tcap > 0 ? align_up(tcap, kW) + 1 : Tw));
  m.d = cv.take<T>(n * R);
  m.v = cv.take<T>(n * R);
  m.t2 = cv.take<T>(n * R);
  m.P = cv.take<T>(tcap > 0 ? 0 : (int64_t)p * nA * R);
  m.lpart = cv.take<T>((int64_t)op->n_partial * R);
  return cv.off + 256;
}

// Launch of the operator kernel in the plan's unit shape (one row per unit, or runs of rows).
template <typename T, int CPL, bool WEIGHTED, bool DOT, typename TW = T>
static void opmul_go(const polar_op* op, const Plan& pl, double shift, const T* X, int64_t ldx, T* Y, int64_t ldy,
                     int R, T* lpart, const CGState& s, int cur, hipStream_t st) {
  if (pl.run > 1)
    hipLaunchKernelGGL((k_opmul<T, CPL, WEIGHTED, DOT, TW, true>), dim3(pl.grid_a), dim3(kBlock), 0, st, *op, shift, X,
                       ldx, Y, ldy, R, pl.L, lpart, s, cur, pl.run);
  else
    hipLaunchKernelGGL((k_opmul<T, CPL, WEIGHTED, DOT, TW, false>), dim3(pl.grid_a), dim3(kBlock), 0, st, *op, shift, X,
                       ldx, Y, ldy, R, pl.L, lpart, s, cur, 1);
}

template <typename T, int CPL>
static int launch_opmul(const polar_op* op, const Plan& pl, double shift, const T* X, int64_t ldx, T* Y,
                        int64_t ldy, int R, T* lpart, const CGState& s, int cur, bool dot, hipStream_t st) {
  if (dot) {
    if (op->weighted)
      opmul_go<T, CPL, true, true>(op, pl, shift, X, ldx, Y, ldy, R, lpart, s, cur, st);
    else
      opmul_go<T, CPL, false, true>(op, pl, shift, X, ldx, Y, ldy, R, lpart, s, cur, st);
  } else {
    if (op->weighted)
      opmul_go<T, CPL, true, false>(op, pl, shift, X, ldx, Y, ldy, R, lpart, s, cur, st);
    else
      opmul_go<T, CPL, false, false>(op, pl, shift, X, ldx, Y, ldy, R, lpart, s, cur, st);
  }
  POLAR_LAUNCH_CHECK();
  if (op->n_long > 0 && pl.grid_l > 0) {
    if (dot)
      hipLaunchKernelGGL((k_longfin<T, CPL, true>), dim3(pl.grid_l), dim3(kBlock), 0, st, *op, shift, X, ldx, Y, ldy, R, pl.L, (const T*)lpart, s, cur, pl.grid_a);
    else
      hipLaunchKernelGGL((k_longfin<T, CPL, false>), dim3(pl.grid_l), dim3(kBlock), 0, st, *op, shift, X, ldx, Y, ldy, R, pl.L, (const T*)lpart, s, cur, pl.grid_a);
    POLAR_LAUNCH_CHECK();
  }
  return 0;
}

// ---------------------------------------------------------------- plain SpMM (drop-in spmm)
template <typename T, int CPL>
static int spmm_impl(const polar_op* op, double shift, const T* X, int64_t ldx, T* Y, int64_t ldy,
                     int64_t R, void* ws, int64_t ws_bytes, hipStream_t st) {
  Plan pl = make_plan(op, R, CPL);
  Carver cv(ws);
  T* lpart = cv.take<T>((int64_t)op->n_partial * R);
  if (cv.off > ws_bytes) return set_error(POLAR_E_ARG, "spmm workspace too small (%lld < %lld)", (long long)ws_bytes, (long long)cv.off);
  CGState s{};
  s.S = 1;
  return launch_opmul<T, CPL>(op, pl, shift, X, ldx, Y, ldy, (int)R, lpart, s, 0, false, st);
}

// ---------------------------------------------------------------- general CSR SpMM (csr_spmm)
// Y[row, :] = sum_k val[k] X[idx[k], :] over the stored entries of each row (diagonal included, any
// values), in entry order, for the column slab blockIdx.y (64 lanes x CPL columns at most).  One
// lane group per row; rows are independent, so nothing needs scratch memory.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_csr_spmm(int64_t n, int R, int L, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ idx, const T* __restrict__ val,
                                                     const T* __restrict__ X, int64_t ldx, T* __restrict__ Y,
                                                     int64_t ldy) {
  Geo g(L);
  const int slab = (int)blockIdx.y * L * CPL;
  const int c0 = slab + g.li * CPL;
  const bool colok = c0 < R;
  const int gpb = kWaves * g.G;
  for (int64_t row = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; row < n; row += (int64_t)gridDim.x * gpb) {
    T acc[CPL];
#pragma unroll
    for (int e = 0; e < CPL; ++e) acc[e] = T(0);
    gather_acc<T, CPL, true>(idx, val, X, ldx, ptr[row], ptr[row + 1], g, colok, c0, acc);
    if (colok) vstore<T, CPL>(Y + row * ldy + c0, acc);
  }
}

template <typename T, int CPL>
static int csr_spmm_impl(int64_t n, int64_t R, const int32_t* ptr, const int32_t* idx, const T* val, const T* X,
                         int64_t ldx, T* Y, int64_t ldy, hipStream_t st) {
  const int L = std::min(64, lanes_for(R, CPL));
  const int64_t gpb = (int64_t)kWaves * (64 / L);
  const int64_t slabs = (R + (int64_t)L * CPL - 1) / ((int64_t)L * CPL);
  if (slabs > 65535) return set_error(POLAR_E_ARG, "R=%lld too wide", (long long)R);
  const int grid = (int)std::min<int64_t>(kOpGridLarge, std::max<int64_t>(1, (n + gpb - 1) / gpb));
  hipLaunchKernelGGL((k_csr_spmm<T, CPL>), dim3(grid, (unsigned)slabs), dim3(kBlock), 0, st, n, (int)R, L, ptr, idx,
                     val, X, ldx, Y, ldy);
  POLAR_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------- msolve
// In-place column compaction of `rows` rows of stride ld: A[row][k] = A[row][perm[k]], k < Rn.
// One wave per row; perm is increasing with perm[k] >= k, every lane's load of block b completes
// before its store of block b, and block b only writes columns below 64(b+1) while later loads
// read columns >= 64(b+1): the in-place move is safe.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_compact_rows(int64_t rows, int64_t ld, int Rn,
                                                         const int* __restrict__ perm, T* A) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < rows; row += nw) {
    T* a = A + row * ld;
    T v[4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (lane + 64 * b < Rn) v[b] = a[perm[lane + 64 * b]];
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (lane + 64 * b < Rn) a[lane + 64 * b] = v[b];
  }
}

// dst[i][k] = src[i][perm[k]] (out of place), i < n, k < Rn.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_compact_vec(int64_t n, int64_t ld, int Rn, const int* __restrict__ perm,
                                                        const T* __restrict__ src, T* __restrict__ dst) {
  const int64_t total = n * Rn;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / Rn;
    const int k = (int)(idx - i * Rn);
    dst[i * ld + k] = src[i * ld + perm[k]];
  }
}

// Write the per-column outputs of the current slots back to the caller's column order:
// h_out/te_out [p][S][Rorig], iters_out [S][Rorig].
__global__ void k_writeback(CGState s, int Rorig, double* __restrict__ h_out, double* __restrict__ te_out,
                            int64_t* __restrict__ iters_out) {
  const int SR = s.S * s.R;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < s.p * SR; idx += gridDim.x * blockDim.x) {
    const int j = idx / SR, rem = idx - j * SR;
    const int g = rem / s.R, c = rem - g * s.R;
    const int col = s.colmap[c];
    const int vc = g * s.ld + c;
    const int64_t o = ((int64_t)j * s.S + g) * Rorig + col;
    if (h_out) h_out[o] = s.h[j * s.sld + vc];
    if (te_out) te_out[o] = s.te[j * s.sld + vc];
    if (j == 0 && iters_out) iters_out[(int64_t)g * Rorig + col] = s.iters[vc];
  }
}

// mask[active[k]] = 1 (the mask is zeroed first)
__global__ void k_mark(const int32_t* __restrict__ active, int64_t nA, uint8_t* __restrict__ mask) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nA; k += (int64_t)gridDim.x * blockDim.x)
    mask[active[k]] = 1;
}

__global__ void k_iota(int* __restrict__ a, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = i;
}

// One CG iteration: rin is the current residual, rout receives the next one (the same vector
// unless the shifted update is blocked; then the k_update launch only updates the direction and
// wt is the iteration's position in its window).
template <typename T, int CPL>
static int run_iter(const polar_op* op, const Plan& pl, double tau0, const T* rin, T* rout, T* d, T* v, T* lpart,
                    T* P, T* Xo, int64_t sj, int64_t si, const int32_t* active, int64_t nA, int p, const CGState& s,
                    int cur, int wt, int hit, hipEvent_t* ev, hipStream_t st) {
  const int64_t n = op->n;
  const int R = s.R;
  const bool longs = op->n_long > 0 && pl.grid_l > 0;
  if (ev) POLAR_HIP(hipEventRecord(ev[0], st));
  if (op->weighted)
    opmul_go<T, CPL, true, true>(op, pl, tau0, (const T*)d, (int64_t)s.ld, v, (int64_t)s.ld, R, lpart, s, cur, st);
  else
    opmul_go<T, CPL, false, true>(op, pl, tau0, (const T*)d, (int64_t)s.ld, v, (int64_t)s.ld, R, lpart, s, cur, st);
  POLAR_LAUNCH_CHECK();
  if (ev) POLAR_HIP(hipEventRecord(ev[1], st));
  if (longs) {
    hipLaunchKernelGGL((k_longfin<T, CPL, true>), dim3(pl.grid_l), dim3(kBlock), 0, st, *op, tau0, (const T*)d, (int64_t)s.ld, v, (int64_t)s.ld, R, pl.L, (const T*)lpart, s, cur, pl.grid_a);
    POLAR_LAUNCH_CHECK();
  }
  if (ev && g_prof_on == 1) POLAR_HIP(hipEventRecord(ev[2], st));
  const int nparts_a = pl.grid_a + (longs ? pl.grid_l : 0);
  hipLaunchKernelGGL(k_alpha, dim3(pl.fin), dim3(kBlock), 0, st, s, nparts_a, cur);
  POLAR_LAUNCH_CHECK();
  if (ev && g_prof_on == 1) POLAR_HIP(hipEventRecord(ev[3], st));
  hipLaunchKernelGGL((k_resid<T, CPL>), dim3(pl.grid_rows), dim3(kBlock), 0, st, n, R, pl.L, rin, rout, (const T*)v,
                     (const T*)d, s, cur);
  POLAR_LAUNCH_CHECK();
  if (ev && g_prof_on == 1) POLAR_HIP(hipEventRecord(ev[4], st));
  hipLaunchKernelGGL(k_scalars, dim3(pl.fin), dim3(kBlock), 0, st, s, pl.grid_rows, cur, wt, hit);
  POLAR_LAUNCH_CHECK();
  if (ev && g_prof_on == 1) POLAR_HIP(hipEventRecord(ev[5], st));
  // blocked: the window flush updates X, P; history: no X, P at all
  const int nb_d = pl.nb_d, nb_s = (s.T > 1 || s.hm) ? 0 : pl.nb_s;
  // few active rows on one segment: spread the poles over copies of the shifted blocks
  const int ps = (s.S <= 1 && nb_s > 0) ? (int)std::min<int64_t>(std::max(p, 1), std::max(1, kSplitGrid / nb_s)) : 1;
  if (p * R <= 1024)
    hipLaunchKernelGGL((k_update<T, CPL, true>), dim3(nb_d + nb_s * ps), dim3(kBlock), 0, st, n, nA, p, R, pl.L,
                       (const T*)rout, d, active, P, Xo, sj, si, nb_d, ps, s, cur);
  else
    hipLaunchKernelGGL((k_update<T, CPL, false>), dim3(nb_d + nb_s * ps), dim3(kBlock), 0, st, n, nA, p, R, pl.L,
                       (const T*)rout, d, active, P, Xo, sj, si, nb_d, ps, s, cur);
  POLAR_LAUNCH_CHECK();
  if (ev && g_prof_on == 1) POLAR_HIP(hipEventRecord(ev[6], st));
  return 0;
}

template <typename T>
static int run_iter_any(int cpl, const polar_op* op, const Plan& pl, double tau0, const T* rin, T* rout, T* d, T* v,
                        T* lpart, T* P, T* Xo, int64_t sj, int64_t si, const int32_t* active, int64_t nA, int p,
                        const CGState& s, int cur, int wt, int hit, hipEvent_t* ev, hipStream_t st) {
  if (cpl == 2) return run_iter<T, 2>(op, pl, tau0, rin, rout, d, v, lpart, P, Xo, sj, si, active, nA, p, s, cur, wt, hit, ev, st);
  if constexpr (sizeof(T) == 4) {
    if (cpl == 4) return run_iter<T, 4>(op, pl, tau0, rin, rout, d, v, lpart, P, Xo, sj, si, active, nA, p, s, cur, wt, hit, ev, st);
  }
  return run_iter<T, 1>(op, pl, tau0, rin, rout, d, v, lpart, P, Xo, sj, si, active, nA, p, s, cur, wt, hit, ev, st);
}

// Residual replacement on the current residual r (after a window's flush): y = (M + tau0) x0 by the
// operator kernel in FP64 on the solve's operator values (long rows in chunks, like every product of
// the solve), then r = w - y on the active columns and their |r|^2.
template <typename T, int C>
static int replace_residual_c(const polar_op* op, const Plan& pl, double tau0, const T* W, int64_t ldw, T* r,
                              double* y64, double* lpart64, const CGState& s, int cur, hipStream_t st) {
  const int R = s.R;
  const double* x0 = s.x0;
  if (op->weighted)
    opmul_go<double, C, true, false, T>(op, pl, tau0, x0, (int64_t)s.ld, y64, (int64_t)s.ld, R, lpart64, s, cur, st);
  else
    opmul_go<double, C, false, false, T>(op, pl, tau0, x0, (int64_t)s.ld, y64, (int64_t)s.ld, R, lpart64, s, cur, st);
  POLAR_LAUNCH_CHECK();
  if (op->n_long > 0 && pl.grid_l > 0) {
    hipLaunchKernelGGL((k_longfin<double, C, false, T>), dim3(pl.grid_l), dim3(kBlock), 0, st, *op, tau0, x0,
                       (int64_t)s.ld, y64, (int64_t)s.ld, R, pl.L, (const double*)lpart64, s, cur, 0);
    POLAR_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_replace_apply<T, C>), dim3(pl.grid_rows), dim3(kBlock), 0, st, op->n, W, ldw,
                     (const double*)y64, r, R, pl.L, s, cur);
  POLAR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_replace_fin, dim3(pl.fin), dim3(kBlock), 0, st, s, pl.grid_rows, cur);
  POLAR_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int replace_residual(int cpl, const polar_op* op, const Plan& pl, double tau0, const T* W, int64_t ldw, T* r,
                            double* y64, double* lpart64, const CGState& s, int cur, hipStream_t st) {
  if (cpl == 2) return replace_residual_c<T, 2>(op, pl, tau0, W, ldw, r, y64, lpart64, s, cur, st);
  if constexpr (sizeof(T) == 4) {
    if (cpl == 4) return replace_residual_c<T, 4>(op, pl, tau0, W, ldw, r, y64, lpart64, s, cur, st);
  }
  return replace_residual_c<T, 1>(op, pl, tau0, W, ldw, r, y64, lpart64, s, cur, st);
}

// End of a window of the blocked update: apply the coefficients to X, P (and, with residual
// replacement, to the seed iterate x0), then start a new window.
template <typename T>
static int flush_window(int cpl, const Plan& pl, const Ring<T>& ring, const int32_t* active, int64_t nA, int p,
                        T* P, T* Xo, int64_t sj, int64_t si, const CGState& s, int64_t n, const T* d, T* dsave,
                        hipStream_t st) {
  if (nA > 0 && pl.nb_s > 0) {
    // column slices of pl.Lf <= 16 lanes (128-byte row pieces): a block's tile has 32 rows, and
    // each staged chunk of poles serves all of them
    const int Lf = pl.Lf;
    const int fc = flush_cpl(cpl);
    const int cw = Lf * fc;
    if (Lf > kFlushLanes || fc * (int)sizeof(T) > 8) return set_error(POLAR_E_ARG, "flush slice too wide");
    const dim3 grid(pl.nb_s, (s.R + cw - 1) / cw);
#define POLAR_FLUSH(C, I) \
  hipLaunchKernelGGL((k_flush<T, C, I>), grid, dim3(kFlushBlock), 0, st, nA, p, s.R, Lf, cw, ring, active, P, Xo, sj, si, s, d, dsave)
    if (fc == 2) {
      if constexpr (sizeof(T) == 4) {
        if (s.ident) POLAR_FLUSH(2, true); else POLAR_FLUSH(2, false);
      }
    } else {
      if (s.ident) POLAR_FLUSH(1, true); else POLAR_FLUSH(1, false);
    }
#undef POLAR_FLUSH
    POLAR_LAUNCH_CHECK();
  }
  if (dsave && (nA < n || nA == 0 || pl.nb_s == 0)) {     // rows the flush did not cover
    if (cpl == 2) {
      hipLaunchKernelGGL((k_seed_flush<T, 2>), dim3(pl.nb_d), dim3(kBlock), 0, st, n, s.R, pl.L, ring, d, dsave, s);
    } else if (cpl == 4) {
      if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((k_seed_flush<T, 4>), dim3(pl.nb_d), dim3(kBlock), 0, st, n, s.R, pl.L, ring, d, dsave, s);
    } else {
      hipLaunchKernelGGL((k_seed_flush<T, 1>), dim3(pl.nb_d), dim3(kBlock), 0, st, n, s.R, pl.L, ring, d, dsave, s);
    }
    POLAR_LAUNCH_CHECK();
  }
  const int64_t tot = (int64_t)p * s.sld;
  hipLaunchKernelGGL(k_coef_reset, dim3((int)std::min<int64_t>(1024, (tot + 255) / 256)), dim3(256), 0, st, s);
  POLAR_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int compact_rows(int64_t rows, int64_t ld, int Rn, const int* perm, T* A, hipStream_t st) {
  const int grid = (int)std::min<int64_t>(65536, std::max<int64_t>(1, (rows * 64 + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_compact_rows<T>, dim3(grid), dim3(kBlock), 0, st, rows, ld, Rn, perm, A);
  POLAR_LAUNCH_CHECK();
  return 0;
}

// Columns per lane for width Rc at row stride ld (both multiples of cpl; 16-byte loads at most).
template <typename T>
static int cpl_for(int64_t Rc, int64_t ld, int cpl_max) {
  for (int c = cpl_max; c > 1; c /= 2)
    if (Rc % c == 0 && ld % c == 0 && (Rc + c - 1) / c <= 64) return c;
  return 1;
}

// Per-thread host resources of the solver, released by polar_release(): the pinned buffer of the
// convergence checks and the stream that hipGraph captures are recorded on.
static thread_local int* t_status = nullptr;
static thread_local int64_t t_status_cap = 0;
static thread_local hipStream_t t_capture = nullptr;
static thread_local int t_capture_dev = -1;

// Pinned host buffer of a thread for the convergence checks (a DMA copy, not a staged one).
static int* status_buffer(int64_t count, int& rc) {
  int*& buf = t_status;
  int64_t& cap = t_status_cap;
  rc = 0;
  if (count > cap) {
    if (buf) (void)hipHostFree(buf);
    buf = nullptr;
    cap = 0;
    void* p = nullptr;
    rc = (int)hipHostMalloc(&p, sizeof(int) * count, hipHostMallocDefault);
    if (rc) return nullptr;
    buf = (int*)p;
    cap = count;
  }
  return buf;
}

// Stream of a thread's captures (default priority; captured work runs on the caller's stream).
static hipStream_t capture_stream(int& rc) {
  hipStream_t& cs = t_capture;
  int& dev = t_capture_dev;
  int cur = 0;
  rc = (int)hipGetDevice(&cur);
  if (rc) return nullptr;
  if (!cs || dev != cur) {
    hipStream_t s = nullptr;
    rc = (int)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (rc) return nullptr;
    if (cs) (void)hipStreamDestroy(cs);      // the previous device's stream
    cs = s;
    dev = cur;
  }
  return cs;
}

// An instantiated hipGraph of one window of CG iterations (parities 0,1,...,1), rebuilt when the plan, the
// width or the vector buffers change.
struct IterGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  void drop() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    x = nullptr;
    g = nullptr;
  }
  ~IterGraph() { drop(); }
};

// Window length of the shifted update: kW with POLAR_BLOCK8 (when the pole list fits the flush
// kernel's LDS), else 1 (per-iteration updates).
static int window_of(int flags, int p) {
  return ((flags & (POLAR_BLOCK8 | POLAR_HISTORY)) && p <= kMaxPolesFlush) ? kW : 1;
}

constexpr int kReplaceEvery = 3 * kW;   // residual replacement period (iterations; whole windows)

template <typename T>
static int msolve_impl(int cpl0, const polar_op* op, const Segs& sg, const T* W, int64_t ldw, int64_t R,
                       const double* tau, int p, const int32_t* active, double tol, int64_t maxit, int flags,
                       T* Xo, int64_t sj, int64_t si, double* h_out, double* te_out, int64_t* iters_out,
                       void* ws, int64_t ws_bytes, hipStream_t st) {
  const int64_t n = op->n;
  const int S = sg.S;
  const int64_t nA = sg.a[S];
  SolveMem<T> m;
  const int Tw = window_of(flags, p);
  const bool repl = (flags & POLAR_REPLACE) != 0;
  const bool hist = (flags & POLAR_HISTORY) != 0;
  if (hist && S != 1) return set_error(POLAR_E_ARG, "POLAR_HISTORY needs one segment");
  if (hist && maxit <= 0) return set_error(POLAR_E_ARG, "POLAR_HISTORY needs maxit >= 1 (the history capacity)");
  const int64_t need = carve_solve<T>(ws, op, S, R, p, nA, Tw, repl, m, hist ? maxit : 0);
  if (need > ws_bytes) return set_error(POLAR_E_ARG, "msolve workspace too small (%lld < %lld)", (long long)ws_bytes, (long long)need);
  CGState s = m.s;
  T *d = m.d, *v = m.v, *t2 = m.t2, *P = m.P, *lpart = m.lpart;
  T* slot[kW];                       // residual ring: r_t of a window in slot[t % Tw]
  for (int t = 0; t < kW; ++t) slot[t] = m.ring + (int64_t)(t % Tw) * n * R;
  int* perm = m.perm;
  s.T = Tw;
  s.R = (int)R;
  s.ld = (int)R;
  s.ident = 1;
  s.p = p;
  s.tol = tol;
  s.freeze = (flags & POLAR_NOFREEZE) ? 0 : 1;
  s.track = (flags & POLAR_TRACK) ? 1 : 0;
  s.trace = (flags & POLAR_TRACE) ? 1 : 0;
  s.S = S;
  s.sld = (int)(S * R);
  s.tau = tau;
  s.tau_dev = 1;
  s.G = sg.G;
  if (sg.G > 1)
    POLAR_HIP(hipMemcpyAsync(s.cgrp, sg.col_group.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
  const int ngroups = std::max(S, sg.G);
  int cpl = cpl0;
  std::vector<char> live(S, 1);
  Plan pl = make_plan(op, R, cpl, &sg, &live, Tw);
  s.nb_a = pl.grid_a;
  if (S > 1) {
    std::vector<int64_t> segh(4 * (S + 1));
    for (int g = 0; g <= S; ++g) {
      segh[g] = sg.row[g];
      segh[(S + 1) + g] = sg.x[g];
      segh[2 * (S + 1) + g] = sg.l[g];
      segh[3 * (S + 1) + g] = sg.a[g];
    }
    POLAR_HIP(hipMemcpyAsync(m.dseg, segh.data(), sizeof(int64_t) * segh.size(), hipMemcpyHostToDevice, st));
    POLAR_HIP(hipMemcpyAsync(m.dtab, pl.tab.data(), sizeof(int) * pl.tab.size(), hipMemcpyHostToDevice, st));
    s.seg_row = m.dseg;
    s.seg_x = m.dseg + (S + 1);
    s.seg_l = m.dseg + 2 * (S + 1);
    s.seg_a = m.dseg + 3 * (S + 1);
    s.bA = m.dtab;
    s.bL = m.dtab + (S + 1);
    s.bR = m.dtab + 2 * (S + 1);
    s.bD = m.dtab + 3 * (S + 1);
    s.bS = m.dtab + 4 * (S + 1);
    if (int rc = zero_async(s.sflag, 2 * S * sizeof(int), st)) return rc;
  }
  const int64_t gpb0 = (int64_t)kWaves * pl.G;

  if (int rc = zero_async(s.flag, 2 * sizeof(int), st)) return rc;
  if (int rc = zero_async(s.stats, 3 * sizeof(unsigned long long), st)) return rc;
  if (repl) {
    if (int rc = zero_async(s.x0, sizeof(double) * (size_t)n * R, st)) return rc;
  }
  if (Tw > 1) {
    // every ring slot holds finite values before a flush reads it (with zero weight or not)
    if (int rc = zero_async(m.ring, sizeof(T) * (size_t)n * R * Tw, st)) return rc;
    hipLaunchKernelGGL(k_coef_reset, dim3((int)std::min<int64_t>(1024, ((int64_t)p * S * R + 255) / 256)), dim3(256), 0, st, s);
    POLAR_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_iota, dim3(1), dim3(256), 0, st, s.colmap, (int)R);
  POLAR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_shift, dim3((p * ngroups + 255) / 256), dim3(256), 0, st, tau, p, ngroups, s.shift);
  POLAR_LAUNCH_CHECK();
  if (cpl == 2) {
    hipLaunchKernelGGL((k_init<T, 2>), dim3(pl.grid_rows), dim3(kBlock), 0, st, n, (int)R, pl.L, W, ldw, slot[0], d, s);
  } else if (cpl == 4) {
    if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((k_init<T, 4>), dim3(pl.grid_rows), dim3(kBlock), 0, st, n, (int)R, pl.L, W, ldw, slot[0], d, s);
  } else {
    hipLaunchKernelGGL((k_init<T, 1>), dim3(pl.grid_rows), dim3(kBlock), 0, st, n, (int)R, pl.L, W, ldw, slot[0], d, s);
  }
  POLAR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_init_fin, dim3(pl.fin), dim3(kBlock), 0, st, s, pl.grid_rows);
  POLAR_LAUNCH_CHECK();
  if (s.amask) {
    if (int rc = zero_async((void*)s.amask, align_up(n, 16), st)) return rc;
    if (nA > 0 && !hist) {     // history: the seed iterate advances on every row in k_seed_flush
      hipLaunchKernelGGL(k_mark, dim3((int)std::min<int64_t>(4096, (nA + 255) / 256)), dim3(256), 0, st, active, nA,
                         (uint8_t*)s.amask);
      POLAR_LAUNCH_CHECK();
    }
  }
  if (m.dsave) POLAR_HIP(hipMemcpyAsync(m.dsave, d, sizeof(T) * (size_t)n * R, hipMemcpyDeviceToDevice, st));
  if (hist) {
    if (int rc = zero_async(s.hist_T, sizeof(int), st)) return rc;
  }
  if (nA > 0 && !hist) {
    const int grid_sh = (int)std::min<int64_t>(65536, std::max<int64_t>(1, (nA + gpb0 - 1) / gpb0));
    if (cpl == 2) {
      hipLaunchKernelGGL((k_init_shift<T, 2>), dim3(grid_sh), dim3(kBlock), 0, st, nA, p, (int)R, pl.L, W, ldw, active, P, Xo, sj, si);
    } else if (cpl == 4) {
      if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((k_init_shift<T, 4>), dim3(grid_sh), dim3(kBlock), 0, st, nA, p, (int)R, pl.L, W, ldw, active, P, Xo, sj, si);
    } else {
      hipLaunchKernelGGL((k_init_shift<T, 1>), dim3(grid_sh), dim3(kBlock), 0, st, nA, p, (int)R, pl.L, W, ldw, active, P, Xo, sj, si);
    }
    POLAR_LAUNCH_CHECK();
  }

  const double tau0 = 0.0;    // unused: the operator kernels read the seed shift from tau (tau_dev)

  const bool prof = g_prof_on != 0;
  const bool use_graph = (flags & POLAR_GRAPH) && !prof && !hist;
  IterGraph graph;
  hipStream_t cs = nullptr;
  if (use_graph) {
    int rc = 0;
    cs = capture_stream(rc);
    if (rc) return hip_fail((hipError_t)rc, "capture stream", __LINE__);
  }
  std::vector<hipEvent_t> ev;    // profiling: kEv events per iteration, for the two batches in flight
  std::vector<float> kms;        // per iteration: per-kernel milliseconds
  std::vector<int> widths;       // per iteration: processed width
  std::vector<char> timed;       // per iteration: its events were recorded (mode 3 samples 1 in 8)
  const int esz = (int)sizeof(T);
  const int align = cpl0;        // compacted widths stay multiples of the initial vector width
  std::vector<int> permh(R);
  // Two batches of kBatch iterations are in flight: batch b+1 is queued before batch b's convergence
  // status is read, so the GPU does not wait for the host at every check.  Launches after
  // convergence do no work (every kernel of an iteration tests the any-active flag first), so the
  // extra batch costs only its launches.  Each batch copies its status into its own half of a
  // pinned buffer: [flag parity 0, flag parity 1, per-column active flags in the batch's slot order].
  const int64_t stat_n = 2 + (int64_t)S * R;
  int* statbuf = nullptr;
  {
    int rc = 0;
    statbuf = status_buffer(2 * stat_n, rc);
    if (rc) return hip_fail((hipError_t)rc, "hipHostMalloc", __LINE__);
  }
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};       // a batch's status copy has landed
    hipEvent_t f[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // profiling of the window flushes
    ~Events() {
      for (auto x : e)
        if (x) (void)hipEventDestroy(x);
      for (auto& y : f)
        for (auto x : y)
          if (x) (void)hipEventDestroy(x);
    }
  } evs;
  for (auto& x : evs.e) POLAR_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  float flush_ms = 0.f;
  auto ring_of = [&]() {
    Ring<T> rg;
    for (int t = 0; t < kW; ++t) rg.s[t] = slot[t];
    return rg;
  };
  int64_t cur_it = -1;         // history: the iteration whose input residual is m.cur (a replacement)
  struct Batch {
    int64_t it0;      // first iteration
    int nit;          // iterations (kBatch, fewer at the cap)
    int buf;          // status half and event set
    int R_at;         // slot width when queued (the layout of its status flags)
    size_t tbase;     // profiling: index of its first iteration in `timed`
  };
  auto launch_batch = [&](const Batch& b) -> int {
    const int64_t it0 = b.it0;
    const int nit = b.nit;
    if (use_graph && nit == kBatch) {
      if (!graph.x) {
        POLAR_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
        int rc = 0;
        for (int k = 0; k < kBatch && !rc; ++k)
          rc = run_iter_any<T>(cpl, op, pl, tau0, slot[k % kW], slot[(k + 1) % kW], d, v, lpart, P, Xo, sj, si, active,
                               nA, p, s, k & 1, k, -1, nullptr, cs);
        if (!rc && Tw > 1) rc = flush_window<T>(cpl, pl, ring_of(), active, nA, p, P, Xo, sj, si, s, n, d, m.dsave, cs);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &gr);
        if (rc) {
          if (gr) (void)hipGraphDestroy(gr);
          return rc;
        }
        POLAR_HIP(ec);
        graph.g = gr;
        POLAR_HIP(hipGraphInstantiate(&graph.x, graph.g, nullptr, nullptr, 0));
      }
      POLAR_HIP(hipGraphLaunch(graph.x, st));
    } else {
      for (int k = 0; k < nit; ++k) {
        const int64_t it = it0 + k;
        hipEvent_t* evp = nullptr;
        if (prof) {
          const int eslot = (b.buf * kBatch + k) * kEv;
          while ((int)ev.size() < eslot + kEv) {
            hipEvent_t e;
            POLAR_HIP(hipEventCreate(&e));
            ev.push_back(e);
          }
          // mode 3: one iteration in every 8, at a position that rotates from one group of 8 to the next
          const bool take = g_prof_on != 3 || (k % 8) == (int)(((it0 + k) / 8) % 8);
          evp = take ? ev.data() + eslot : nullptr;
          widths.push_back(s.R);
          timed.push_back(take ? 1 : 0);
        }
        const T* rin = slot[k % kW];
        T* rout = slot[(k + 1) % kW];
        if (hist) {       // r_t in ring slot t (or the residual that replaced it), r_{t+1} into slot t+1
          rin = it == cur_it ? m.cur : m.ring + it * n * R;
          rout = m.ring + (it + 1) * n * R;
        }
        int rc = run_iter_any<T>(cpl, op, pl, tau0, rin, rout, d, v, lpart, P, Xo, sj, si, active, nA, p, s,
                                 (int)(it & 1), k, hist ? (int)it : -1, evp, st);
        if (rc) return rc;
      }
      if (Tw > 1) {
        const bool fprof = prof && g_prof_on == 1;
        if (fprof) {
          for (auto& e : evs.f[b.buf])
            if (!e) POLAR_HIP(hipEventCreate(&e));
          POLAR_HIP(hipEventRecord(evs.f[b.buf][0], st));
        }
        if (hist) {
          // no shifted blocks: only the seed iterate (on every row) and the next window's start
          Ring<T> rg;
          for (int t = 1; t <= kW; ++t) rg.s[t % kW] = m.ring + (it0 + t) * n * R;
          if (int rc = flush_window<T>(cpl, pl, rg, active, 0, p, P, Xo, sj, si, s, n, d, m.dsave, st)) return rc;
        } else {
          if (int rc = flush_window<T>(cpl, pl, ring_of(), active, nA, p, P, Xo, sj, si, s, n, d, m.dsave, st)) return rc;
        }
        if (fprof) POLAR_HIP(hipEventRecord(evs.f[b.buf][1], st));
      }
    }
    if (repl && nit == kBatch && (it0 + kBatch) % kReplaceEvery == 0) {
      // the current residual is slot[0] after a full window (Tw = kW) or the only vector (Tw = 1);
      // after convergence the replacement kernels return at once, like the iterations.  History:
      // ring slot t keeps the residual the shifted recurrences used, the replaced one goes to `cur`
      T* rcur = slot[0];
      if (hist) {
        const int64_t t1 = it0 + kBatch;
        POLAR_HIP(hipMemcpyAsync(m.cur, m.ring + t1 * n * R, sizeof(T) * (size_t)n * R, hipMemcpyDeviceToDevice, st));
        rcur = m.cur;
        cur_it = t1;
      }
      if (int rc = replace_residual<T>(cpl, op, pl, tau0, W, ldw, rcur, m.y64, m.lpart64, s,
                                       (int)((it0 + kBatch) & 1), st))
        return rc;
    }
    POLAR_HIP(hipMemcpyAsync(statbuf + b.buf * stat_n, s.flag, sizeof(int) * (2 + (S > 1 ? S * R : s.R)),
                             hipMemcpyDeviceToHost, st));
    POLAR_HIP(hipEventRecord(evs.e[b.buf], st));
    return 0;
  };
  bool converged = false;
  bool inflight = false;       // a queued batch whose status has not been read
  int inflight_buf = 0;
  if (maxit > 0) {
    Batch cur{0, (int)std::min<int64_t>(kBatch, maxit), 0, s.R, timed.size()};
    if (int rc = launch_batch(cur)) return rc;
    for (;;) {
      const int64_t it1 = cur.it0 + cur.nit;
      const bool have_next = it1 < maxit;
      Batch nxt{it1, (int)std::min<int64_t>(kBatch, maxit - it1), cur.buf ^ 1, s.R, timed.size()};
      if (have_next) {
        if (int rc = launch_batch(nxt)) return rc;
      }
      inflight = have_next;
      inflight_buf = nxt.buf;
      POLAR_HIP(hipEventSynchronize(evs.e[cur.buf]));
      const int* stat = statbuf + cur.buf * stat_n;
      const int flag = stat[(cur.it0 + cur.nit) & 1];
      if (prof) {
        for (int k = 0; k < cur.nit; ++k)
          for (int c = 0; c < kEv - 1; ++c) {
            float ms = 0.f;
            const int eslot = (cur.buf * kBatch + k) * kEv;
            if (timed[cur.tbase + k] && (c == 0 || g_prof_on == 1))
              POLAR_HIP(hipEventElapsedTime(&ms, ev[eslot + c], ev[eslot + c + 1]));
            kms.push_back(ms);
          }
        if (Tw > 1 && g_prof_on == 1 && evs.f[cur.buf][0]) {
          float ms = 0.f;
          POLAR_HIP(hipEventElapsedTime(&ms, evs.f[cur.buf][0], evs.f[cur.buf][1]));
          flush_ms += ms;
        }
      }
      if (!flag) {
        converged = true;
        break;
      }
      if (!have_next) break;
      // Re-planning and compaction follow the per-column flags of `cur` (a superset of the columns
      // still active after `nxt`), in the slot order `cur` was queued with; they are queued after
      // `nxt`, which must be a whole batch (the current residual is then slot[0]).
      const int* acth = stat + 2;
      const bool may_change = cur.R_at == s.R && nxt.nit == kBatch;
      if (may_change && S > 1) {
        // Re-plan over the segments still iterating: finished segments get no blocks, the others
        // share the whole budget (their kernels skip finished segments either way).
        bool changed = false;
        for (int g = 0; g < S; ++g) {
          char on = 0;
          for (int c = 0; c < R; ++c) on |= acth[g * R + c] != 0;
          changed |= on != live[g];
          live[g] = on;
        }
        if (changed) {
          if (use_graph) POLAR_HIP(hipStreamSynchronize(st));   // the batch in flight may be a graph launch
          pl = make_plan(op, R, cpl, &sg, &live, Tw);
          s.nb_a = pl.grid_a;
          POLAR_HIP(hipMemcpyAsync(m.dtab, pl.tab.data(), sizeof(int) * pl.tab.size(), hipMemcpyHostToDevice, st));
          graph.drop();
        }
      } else if (may_change && !hist) {
        // Active-column compaction (one segment): when at most half of the current slots are still
        // active, move them (padded to the vector width with finished slots) to the front.  Columns
        // are independent, so every column's arithmetic is unchanged.
        int count = 0;
        for (int c = 0; c < s.R; ++c) count += acth[c] != 0;
        const int Rn = std::min(s.R, (int)align_up(std::max(count, 1), align));
        if (Rn <= s.R / 2) {
          if (use_graph) POLAR_HIP(hipStreamSynchronize(st));   // the batch in flight may be a graph launch
          int k = 0;
          for (int c = 0; c < s.R; ++c)
            if (acth[c]) permh[k++] = c;
          for (int c = 0; c < s.R && k < Rn; ++c)
            if (!acth[c]) permh[k++] = c;
          POLAR_HIP(hipMemcpyAsync(perm, permh.data(), sizeof(int) * Rn, hipMemcpyHostToDevice, st));
          hipLaunchKernelGGL(k_writeback, dim3((p * s.R + 255) / 256), dim3(256), 0, st, s, (int)R, h_out, te_out, iters_out);
          POLAR_LAUNCH_CHECK();
          // the current residual is slot[0] (a full window ends with r_8 there; Tw = 1: the residual)
          const int gv = (int)std::min<int64_t>(65536, std::max<int64_t>(1, (n * Rn + kBlock - 1) / kBlock));
          hipLaunchKernelGGL(k_compact_vec<T>, dim3(gv), dim3(kBlock), 0, st, n, (int64_t)R, Rn, (const int*)perm, (const T*)slot[0], v);
          POLAR_LAUNCH_CHECK();
          hipLaunchKernelGGL(k_compact_vec<T>, dim3(gv), dim3(kBlock), 0, st, n, (int64_t)R, Rn, (const int*)perm, (const T*)d, t2);
          POLAR_LAUNCH_CHECK();
          {
            T* old0 = slot[0];
            for (int t = 0; t < kW; ++t)
              if (slot[t] == old0) slot[t] = v;
            v = old0;
          }
          std::swap(d, t2);
          if (repl)
            if (int e = compact_rows<double>(n, R, Rn, perm, s.x0, st)) return e;
          if (m.dsave)
            if (int e = compact_rows<T>(n, R, Rn, perm, m.dsave, st)) return e;
          if (int e = compact_rows<T>((int64_t)p * nA, R, Rn, perm, P, st)) return e;
          for (double* a : {s.rr, s.stop, s.a, s.be, s.ap, s.bp})
            if (int e = compact_rows<double>(1, R, Rn, perm, a, st)) return e;
          for (int* a : {s.act, s.dmask})
            if (int e = compact_rows<int>(1, R, Rn, perm, a, st)) return e;
          if (int e = compact_rows<int64_t>(1, R, Rn, perm, s.iters, st)) return e;
          for (double* a : {s.z, s.zp, s.zn, s.aj, s.bj, s.h, s.ip, s.pn, s.te, s.tg})
            if (int e = compact_rows<double>(p, R, Rn, perm, a, st)) return e;
          for (int* a : {s.live, s.m})
            if (int e = compact_rows<int>(p, R, Rn, perm, a, st)) return e;
          if (int e = compact_rows<int>(1, R, Rn, perm, s.colmap, st)) return e;
          if (sg.G > 1)
            if (int e = compact_rows<int>(1, R, Rn, perm, s.cgrp, st)) return e;
          s.R = Rn;
          s.ident = 0;
          cpl = cpl_for<T>(Rn, R, cpl0);
          pl = make_plan(op, Rn, cpl, &sg, &live, Tw);
          s.nb_a = pl.grid_a;
          graph.drop();
        }
      }
      cur = nxt;
    }
  }
  // the batch queued after the last check (it does no work after convergence) has landed its
  // status before the pinned buffer is used again
  if (inflight) POLAR_HIP(hipEventSynchronize(evs.e[inflight_buf]));
  if (prof) {
    // count only the iterations that did work (launches after convergence return at once)
    std::vector<int64_t> ih((size_t)S * s.R);
    unsigned long long st2[3];
    if (S == 1) {
      POLAR_HIP(hipMemcpyAsync(ih.data(), s.iters, sizeof(int64_t) * s.R, hipMemcpyDeviceToHost, st));
    } else {
      POLAR_HIP(hipMemcpyAsync(ih.data(), s.iters, sizeof(int64_t) * S * R, hipMemcpyDeviceToHost, st));
    }
    POLAR_HIP(hipMemcpyAsync(st2, s.stats, sizeof(st2), hipMemcpyDeviceToHost, st));
    POLAR_HIP(hipStreamSynchronize(st));
    int64_t real = 0;
    for (int64_t v2 : ih) real = std::max(real, v2);
    real = std::min<int64_t>(real, (int64_t)kms.size() / (kEv - 1));
    int32_t nnz_h = 0;
    POLAR_HIP(hipMemcpy(&nnz_h, op->indptr + op->n, sizeof(int32_t), hipMemcpyDeviceToHost));
    int64_t sampled = 0;
    for (int64_t k = 0; k < real; ++k) {
      if (!timed[k]) continue;
      ++sampled;
      for (int c = 0; c < kEv - 1; ++c) g_prof[c].ms += kms[(kEv - 1) * k + c];
      const double w = (double)widths[k] * S;
      g_prof[0].bytes += 4.0 * (n + 1) + 4.0 * nnz_h + (op->weighted ? (double)esz * nnz_h : 0.0) +
                         (double)esz * n + 2.0 * esz * (double)n * widths[k];
      g_prof[1].bytes += (double)esz * widths[k] * ((double)op->n_partial + 2.0 * op->n_long);
      g_prof[2].bytes += 16.0 * widths[k] * (pl.grid_a + pl.grid_l);
      g_prof[4].bytes += 16.0 * widths[k] * pl.grid_rows + 120.0 * w * p;
    }
    for (int c = 0; c < kEv - 1; ++c) g_prof[c].launches += sampled;
    g_prof[kEv].launches += real;
    // active columns summed over iterations; with segments a column-iteration covers its
    // segment's rows only, so the row streams are counted at the mean segment length
    const double nrows = (double)n / S, narows = (double)nA / S;
    g_prof[3].bytes += 3.0 * esz * nrows * (double)st2[0];
    if (Tw > 1) {
      // direction update per active column-iteration; per window the flush reads the T residuals
      // of every column slot and reads and writes X, P of the (pole, column) pairs it changed
      double slots = 0.0;
      for (int64_t k = 0; k < (int64_t)(kms.size() / (kEv - 1)) && k < (int64_t)widths.size(); ++k)
        slots += (double)widths[k] * S;     // iterations whose status was read
      g_prof[5].bytes += esz * 3.0 * nrows * (double)st2[0] + esz * narows * slots + 4.0 * esz * narows * (double)st2[2];
      g_prof[5].ms += flush_ms;
    } else {
      g_prof[5].bytes += esz * (3.0 * nrows + narows) * (double)st2[0] + 4.0 * esz * narows * (double)st2[1];
    }
    g_prof[6].launches += 1;
    g_prof[6].bytes += (double)st2[1];
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  hipLaunchKernelGGL(k_writeback, dim3((p * S * s.R + 255) / 256), dim3(256), 0, st, s, (int)R, h_out, te_out, iters_out);
  POLAR_LAUNCH_CHECK();
  if (!converged) return set_error(POLAR_E_ITERCAP, "multi-shift CG iteration cap");
  return 0;
}

template <typename T>
static int64_t msolve_ws(const polar_op* op, int S, int64_t R, int p, int64_t nA, int flags) {
  SolveMem<T> m;
  return carve_solve<T>(nullptr, op, S, R, p, nA, window_of(flags, p), (flags & POLAR_REPLACE) != 0, m);
}

// Segment ranges from the caller's table, checked against the operator.
static int segs_from(const polar_op* op, const polar_segs* in, int64_t R, Segs& sg) {
  if (!in || in->S < 1 || !in->rows || !in->extra || !in->longs || !in->active)
    return set_error(POLAR_E_ARG, "invalid polar_segs");
  const int S = in->S;
  sg.S = S;
  sg.row.assign(in->rows, in->rows + S + 1);
  sg.x.assign(in->extra, in->extra + S + 1);
  sg.l.assign(in->longs, in->longs + S + 1);
  sg.a.assign(in->active, in->active + S + 1);
  if (sg.row[0] != 0 || sg.row[S] != op->n || sg.x[0] != 0 || sg.x[S] != op->n_extra || sg.l[0] != 0 ||
      sg.l[S] != op->n_long || sg.a[0] != 0)
    return set_error(POLAR_E_ARG, "segment ranges do not cover the operator");
  for (int g = 0; g < S; ++g)
    if (sg.row[g + 1] <= sg.row[g] || sg.x[g + 1] < sg.x[g] || sg.l[g + 1] < sg.l[g] || sg.a[g + 1] < sg.a[g])
      return set_error(POLAR_E_ARG, "segment %d: empty or decreasing range", g);
  sg.G = in->G > 1 ? in->G : 1;
  if (sg.G > 1) {
    if (S != 1) return set_error(POLAR_E_ARG, "column groups need a single segment");
    if (!in->col_group) return set_error(POLAR_E_ARG, "column groups without col_group");
    sg.col_group.assign(in->col_group, in->col_group + R);
    for (int v : sg.col_group)
      if (v < 0 || v >= sg.G) return set_error(POLAR_E_ARG, "column group %d out of range", v);
  }
  return 0;
}

static Segs one_seg(const polar_op* op, int64_t nA) {
  Segs sg;
  sg.S = 1;
  sg.row = {0, op->n};
  sg.x = {0, op->n_extra};
  sg.l = {0, op->n_long};
  sg.a = {0, nA};
  return sg;
}

template <typename T>
static int msolve_dispatch(const polar_op* op, const Segs& sg, const void* W, int64_t ldw, int64_t R,
                           const double* tau, int32_t p, const int32_t* active, double tol, int64_t maxit,
                           int32_t flags, void* Xout, int64_t x_sj, int64_t x_si, double* h_out,
                           double* te_out, int64_t* iters_out, void* ws, int64_t ws_bytes, hipStream_t st) {
  const int cpl = pick_cpl(op->dtype, R, {{W, ldw}, {Xout, x_si}, {(const char*)Xout + 0, x_sj}}, op->cpl_max);
  if (lanes_for(R, cpl) > 64) return set_error(POLAR_E_ARG, "R=%lld too wide for one wave at this alignment", (long long)R);
  return msolve_impl<T>(cpl, op, sg, (const T*)W, ldw, R, tau, p, active, tol, maxit, flags, (T*)Xout, x_sj, x_si,
                        h_out, te_out, iters_out, ws, ws_bytes, st);
}

// ---------------------------------------------------------------- zero-allocation SpMM
// polar_spmm_static keeps the long-row chunk partials in this file-scope device array instead of a
// caller workspace.  Every GPU the library runs on holds its own copy until the process ends.  In
// host code the array's name is not a device address: its address on the current device is looked
// up once per device (hipGetSymbolAddress, with that device current) and only that address is
// passed to kernels.  Calls that overlap on one GPU share the array and race.
constexpr int64_t kScratchBytes = int64_t(32) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_address(void** out) {
  int dev = 0;
  POLAR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(POLAR_E_ARG, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    POLAR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(POLAR_E_ARG, "no device address for the static scratch on device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// ------------------------------------------------------------------------------------------------
// POLAR_HISTORY: X_j formed after the solve from the stored residuals r_0 .. r_T and the per-step
// pole coefficients.  With P^(0) = r_0 and, at every step t where the pole is live,
// X += a_t P^(t), P^(t+1) = z_t r_{t+1} + b_t P^(t) (no change otherwise), X_j = sum_t w_t r_t with
// weights from the backward recursion G_T = 0; at a live step G_t = a_t + b_t G_{t+1} and
// w_{t+1} = z_t G_{t+1}, at any other step G_t = G_{t+1} and w_{t+1} = 0; w_0 = G_0.  FP64 weights
// and sums, whatever the vector type.  Loops over steps load kHistBatch steps before using them,
// so that their latencies overlap.
constexpr int kHistBatch = 16;

// One wave per (selected column i, pole j).  Lane l takes the steps [t0, t1) of the l-th of 64 equal
// segments: pass 1 composes the segment's map G_{t0} = A + B G_{t1} (the identity at frozen steps),
// a backward scan over the lanes composes the maps of all later segments, which gives every
// segment's G_{t1} (G_T = 0 after the last), and pass 2 runs the recursion inside the segment from
// there and writes the segment's weights.
__global__ __launch_bounds__(256) void k_hist_coef(const uint8_t* __restrict__ hm, const double* __restrict__ haj,
                                                   const double* __restrict__ hbj, const double* __restrict__ hzn,
                                                   const int* __restrict__ hist_T, const int64_t* __restrict__ iters,
                                                   int p, int ld, int64_t tcap, const int* __restrict__ cols, int C,
                                                   double* __restrict__ coef) {
  const int lane = threadIdx.x & 63;
  const int pair = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pair >= C * p) return;               // the whole wave
  const int i = pair / p, j = pair - i * p;
  const int c = cols ? cols[i] : i;
  // a column is active in its first iters[c] steps only (it never becomes active again), so its
  // weights w[t] vanish for t > iters[c]: only w[0 .. iters[c]] are formed and read
  const int Tn = (int)min((int64_t)hist_T[0], iters[c]);
  double* w = coef + ((int64_t)c * p + j) * (tcap + 1);
  const int seg = (Tn + 63) / 64;
  const int t0 = min(Tn, lane * seg), t1 = min(Tn, t0 + seg);
  double A = 0.0, B = 1.0;
  for (int tb = t1 - 1; tb >= t0; tb -= kHistBatch) {
    const int nb = min(kHistBatch, tb - t0 + 1);
    uint8_t mm[kHistBatch];
    double aa[kHistBatch], bb[kHistBatch];
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u)
      if (u < nb) {
        const int64_t hx = ((int64_t)(tb - u) * p + j) * ld + c;
        mm[u] = hm[hx];
        aa[u] = haj[hx];
        bb[u] = hbj[hx];
      }
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u)
      if (u < nb && mm[u]) {
        A = aa[u] + bb[u] * A;
        B = bb[u] * B;
      }
  }
  // lane l: f_l o f_{l+1} o ... o f_63
  for (int off = 1; off < 64; off <<= 1) {
    const double A2 = __shfl_down(A, off, 64), B2 = __shfl_down(B, off, 64);
    if (lane + off < 64) {
      A = A + B * A2;
      B = B * B2;
    }
  }
  double G = __shfl_down(A, 1, 64);
  if (lane == 63) G = 0.0;
  for (int tb = t1 - 1; tb >= t0; tb -= kHistBatch) {
    const int nb = min(kHistBatch, tb - t0 + 1);
    uint8_t mm[kHistBatch];
    double aa[kHistBatch], bb[kHistBatch], zz[kHistBatch];
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u)
      if (u < nb) {
        const int64_t hx = ((int64_t)(tb - u) * p + j) * ld + c;
        mm[u] = hm[hx];
        aa[u] = haj[hx];
        bb[u] = hbj[hx];
        zz[u] = hzn[hx];
      }
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u)
      if (u < nb) {
        if (mm[u]) {
          w[tb - u + 1] = zz[u] * G;
          G = aa[u] + bb[u] * G;
        } else {
          w[tb - u + 1] = 0.0;
        }
      }
  }
  if (lane == 0) w[0] = G;
}

// E[(j*K + k)*R + c] = X_j[rows[k*R + c], c]: one wave per value, lane l summing the steps
// t = l (mod 64), the lanes' FP64 sums added across the wave.
template <typename T>
__global__ __launch_bounds__(256) void k_hist_rows(const T* __restrict__ ring, int64_t n, int ld, int p, int64_t tcap,
                                                   const int* __restrict__ hist_T, const int64_t* __restrict__ iters,
                                                   const double* __restrict__ coef, int R, int K,
                                                   const int32_t* __restrict__ rows, T* __restrict__ E) {
  constexpr int kU = 4;
  const int lane = threadIdx.x & 63;
  const int idx = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (idx >= p * K * R) return;            // the whole wave
  const int c = idx % R, k = (idx / R) % K, j = idx / (R * K);
  const int64_t row = rows[k * R + c];
  const int Tn = (int)min((int64_t)hist_T[0], iters[c]);     // the column's last nonzero weight
  const double* w = coef + ((int64_t)c * p + j) * (tcap + 1);
  double acc = 0.0;
  for (int tb = lane; tb <= Tn; tb += 64 * kU) {
    double wt[kU], v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int t = tb + 64 * u;
      wt[u] = 0.0;
      v[u] = 0.0;
      if (t <= Tn) {
        wt[u] = w[t];
        v[u] = (double)ring[(int64_t)t * n * ld + row * ld + c];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (wt[u] != 0.0) acc = acc + wt[u] * v[u];
  }
  for (int off = 32; off > 0; off >>= 1) acc = acc + __shfl_xor(acc, off, 64);
  if (lane == 0) E[idx] = (T)acc;
}

// X[(i*p + j)*nA + a] = X_j[active[a], cols[i]]: one thread per (active row, selected column), the C
// columns of a row on neighbouring lanes (one residual row serves them all); the weights are staged
// in LDS kHistStage steps at a time; PM >= p accumulators per thread.  Each column's weight block
// is padded by kHistPad doubles, so that the C lanes of a row, which read the same step and pole of
// different columns at once, hit different LDS banks.
constexpr int kHistStage = 32;
constexpr int kHistMaxC = 4;
constexpr int kHistPad = 2;
template <typename T, int PM>
__global__ __launch_bounds__(256) void k_hist_cols(const T* __restrict__ ring, int64_t n, int ld, int p, int64_t tcap,
                                                   const int* __restrict__ hist_T, const int64_t* __restrict__ iters,
                                                   const double* __restrict__ coef, const int* __restrict__ cols,
                                                   int C, const int32_t* __restrict__ active, int64_t nA,
                                                   T* __restrict__ X) {
  __shared__ double sw_[kHistMaxC][PM * kHistStage + kHistPad];
  auto sw = [&](int ci, int j, int tt) -> double& { return sw_[ci][j * kHistStage + tt]; };
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(q % C);
  const int64_t a = q / C;
  const int c = cols[i];
  // each column's weights end at its iteration count (k_hist_coef); the block walks to the largest
  int Tc[kHistMaxC];
  int Te = 0;
#pragma unroll
  for (int ci = 0; ci < kHistMaxC; ++ci) {
    Tc[ci] = ci < C ? (int)min((int64_t)hist_T[0], iters[cols[ci]]) : 0;
    Te = max(Te, Tc[ci]);
  }
  // this lane's column (selected without a dynamic index, so that Tc stays in registers)
  static_assert(kHistMaxC == 4, "the selection below covers 4 columns");
  const int Ti = Tc[0] * (i == 0) + Tc[1] * (i == 1) + Tc[2] * (i == 2) + Tc[3] * (i == 3);
  double acc[PM];
#pragma unroll
  for (int j = 0; j < PM; ++j) acc[j] = 0.0;
  const bool ok = a < nA;
  const int64_t row = ok ? active[a] : 0;
  for (int t0 = 0; t0 <= Te; t0 += kHistStage) {
    __syncthreads();
    for (int e = threadIdx.x; e < C * p * kHistStage; e += blockDim.x) {
      const int ci = e / (p * kHistStage), j = (e / kHistStage) % p, tt = e % kHistStage;
      const int tcol = (int)min((int64_t)hist_T[0], iters[cols[ci]]);
      sw(ci, j, tt) = t0 + tt <= tcol ? coef[((int64_t)cols[ci] * p + j) * (tcap + 1) + t0 + tt] : 0.0;
    }
    __syncthreads();
    if (ok) {
      const int nt = min(kHistStage, Ti + 1 - t0);
      for (int u0 = 0; u0 < nt; u0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (u0 + u < nt)       // the history is read once: non-temporal loads
            v[u] = (double)__builtin_nontemporal_load(ring + (int64_t)(t0 + u0 + u) * n * ld + row * ld + c);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (u0 + u < nt) {
#pragma unroll
            for (int j = 0; j < PM; ++j)
              if (j < p) acc[j] = acc[j] + sw(i, j, u0 + u) * v[u];
          }
      }
    }
  }
  if (ok) {
#pragma unroll
    for (int j = 0; j < PM; ++j)
      if (j < p) X[((int64_t)i * p + j) * nA + a] = (T)acc[j];
  }
}

template <typename T>
static int history_mem(const polar_op* op, int64_t R, int p, int64_t nA, int flags, int64_t tcap, void* ws,
                       int64_t ws_bytes, SolveMem<T>& m) {
  if (!(flags & POLAR_HISTORY)) return set_error(POLAR_E_ARG, "history query without POLAR_HISTORY");
  if (R <= 0 || p <= 0 || nA <= 0 || tcap <= 0 || !ws) return set_error(POLAR_E_ARG, "bad history arguments");
  const int64_t need = carve_solve<T>(ws, op, 1, R, p, nA, window_of(flags, p), (flags & POLAR_REPLACE) != 0, m, tcap);
  if (need > ws_bytes) return set_error(POLAR_E_ARG, "history workspace too small (%lld < %lld)", (long long)ws_bytes, (long long)need);
  return 0;
}

template <typename T>
static int history_rows_impl(const polar_op* op, int64_t R, int p, int64_t nA, int flags, int64_t tcap, void* ws,
                             int64_t ws_bytes, int K, const int32_t* rows, T* E, hipStream_t st) {
  SolveMem<T> m;
  if (int rc = history_mem<T>(op, R, p, nA, flags, tcap, ws, ws_bytes, m)) return rc;
  const int64_t tot = (int64_t)p * K * R;
  if (tot > INT32_MAX || R * p > INT32_MAX) return set_error(POLAR_E_ARG, "too many history values");
  const CGState& s = m.s;
  // one wave per (column, pole) pair, then one wave per value: 4 waves per block
  hipLaunchKernelGGL(k_hist_coef, dim3((int)((R * p + 3) / 4)), dim3(256), 0, st, s.hm, s.haj, s.hbj, s.hzn,
                     s.hist_T, (const int64_t*)s.iters, p, (int)R, tcap, (const int*)nullptr, (int)R, m.coef);
  POLAR_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_hist_rows<T>), dim3((int)((tot + 3) / 4)), dim3(256), 0, st, (const T*)m.ring, op->n,
                     (int)R, p, tcap, s.hist_T, (const int64_t*)s.iters, (const double*)m.coef, (int)R, K, rows, E);
  POLAR_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int history_cols_impl(const polar_op* op, int64_t R, int p, int64_t nA, int flags, int64_t tcap, void* ws,
                             int64_t ws_bytes, int C, const int32_t* cols, const int32_t* active, T* X,
                             hipStream_t st) {
  SolveMem<T> m;
  if (int rc = history_mem<T>(op, R, p, nA, flags, tcap, ws, ws_bytes, m)) return rc;
  if (p > 64) return set_error(POLAR_E_ARG, "polar_history_cols supports at most 64 poles (p=%d)", p);
  if (C > kHistMaxC) return set_error(POLAR_E_ARG, "polar_history_cols takes at most %d columns per call", kHistMaxC);
  const int64_t blocks = (nA * C + 255) / 256;
  if (blocks > INT32_MAX) return set_error(POLAR_E_ARG, "too many rows x columns");
  const CGState& s = m.s;
  hipLaunchKernelGGL(k_hist_coef, dim3((C * p + 3) / 4), dim3(256), 0, st, s.hm, s.haj, s.hbj, s.hzn, s.hist_T,
                     (const int64_t*)s.iters, p, (int)R, tcap, cols, C, m.coef);
  POLAR_LAUNCH_CHECK();
  const dim3 grid((unsigned)blocks);
  if (p <= 32)
    hipLaunchKernelGGL((k_hist_cols<T, 32>), grid, dim3(256), 0, st, (const T*)m.ring, op->n, (int)R, p, tcap,
                       s.hist_T, (const int64_t*)s.iters, (const double*)m.coef, cols, C, active, nA, X);
  else
    hipLaunchKernelGGL((k_hist_cols<T, 64>), grid, dim3(256), 0, st, (const T*)m.ring, op->n, (int)R, p, tcap,
                       s.hist_T, (const int64_t*)s.iters, (const double*)m.coef, cols, C, active, nA, X);
  POLAR_LAUNCH_CHECK();
  return 0;
}

}  // namespace polar

using namespace polar;

extern "C" {

int polar_abi_version(void) { return POLAR_ABI_VERSION; }

int polar_prof_set(int32_t on) {
  if (on < 0 || on > 3) return set_error(POLAR_E_ARG, "profiler mode %d (0..3)", on);
  g_prof_on = on;
  for (auto& c : g_prof) c = ProfClass();
  return 0;
}

int polar_prof_get(int32_t cls, double* ms, int64_t* launches, double* bytes) {
  if (cls < 0 || cls >= kProfClasses) return set_error(POLAR_E_ARG, "bad profile class");
  if (ms) *ms = g_prof[cls].ms;
  if (launches) *launches = g_prof[cls].launches;
  if (bytes) *bytes = g_prof[cls].bytes;
  return 0;
}

const char* polar_last_error(void) { return g_err; }

int polar_release(void) {
  int rc = 0;
  if (t_status) {
    const hipError_t e = hipHostFree(t_status);
    if (e != hipSuccess) rc = hip_fail(e, "hipHostFree", __LINE__);
    t_status = nullptr;
    t_status_cap = 0;
  }
  if (t_capture) {
    const hipError_t e = hipStreamDestroy(t_capture);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamDestroy", __LINE__);
    t_capture = nullptr;
    t_capture_dev = -1;
  }
  return rc;
}

int64_t polar_spmm_ws_bytes(const polar_op* op, int64_t R) {
  if (check_op(op)) return -1;
  return align_up((int64_t)op->n_partial * R * elem_size(op->dtype), 256) + 512;
}

int polar_spmm(const polar_op* op, double shift, const void* X, int64_t ldx, void* Y, int64_t ldy,
               int64_t R, void* ws, int64_t ws_bytes, void* stream) {
  if (int rc = check_op(op)) return rc;
  if (R <= 0 || ldx < R || ldy < R) return set_error(POLAR_E_ARG, "bad R/ld");
  hipStream_t st = (hipStream_t)stream;
  const int cpl = pick_cpl(op->dtype, R, {{X, ldx}, {Y, ldy}, {ws, 0}}, op->cpl_max);
  if (lanes_for(R, cpl) > 64) return set_error(POLAR_E_ARG, "R=%lld too wide for one wave (max %d)", (long long)R, 64 * cpl);
  if (op->dtype == POLAR_F64) {
    if (cpl == 2) return spmm_impl<double, 2>(op, shift, (const double*)X, ldx, (double*)Y, ldy, R, ws, ws_bytes, st);
    return spmm_impl<double, 1>(op, shift, (const double*)X, ldx, (double*)Y, ldy, R, ws, ws_bytes, st);
  }
  if (cpl == 4) return spmm_impl<float, 4>(op, shift, (const float*)X, ldx, (float*)Y, ldy, R, ws, ws_bytes, st);
  if (cpl == 2) return spmm_impl<float, 2>(op, shift, (const float*)X, ldx, (float*)Y, ldy, R, ws, ws_bytes, st);
  return spmm_impl<float, 1>(op, shift, (const float*)X, ldx, (float*)Y, ldy, R, ws, ws_bytes, st);
}

int64_t polar_spmm_static_max_bytes(void) { return kScratchBytes; }

int polar_spmm_static(const polar_op* op, double shift, const void* X, int64_t ldx, void* Y, int64_t ldy,
                      int64_t R, void* stream) {
  if (int rc = check_op(op)) return rc;
  if (R <= 0 || ldx < R || ldy < R) return set_error(POLAR_E_ARG, "bad R/ld");
  const int64_t need = polar_spmm_ws_bytes(op, R);
  if (need > kScratchBytes)
    return set_error(POLAR_E_ARG,
                     "polar_spmm_static: %lld scratch bytes needed (%d long-row chunks x R=%lld) exceed the static "
                     "maximum of %lld; call polar_spmm with a workspace",
                     (long long)need, op->n_partial, (long long)R, (long long)kScratchBytes);
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  return polar_spmm(op, shift, X, ldx, Y, ldy, R, ws, kScratchBytes, stream);
}

int polar_csr_spmm(int32_t dtype, int64_t n, int64_t R, const int32_t* indptr, const int32_t* indices,
                   const void* values, const void* X, int64_t ldx, void* Y, int64_t ldy, void* stream) {
  if (!dtype_ok(dtype)) return set_error(POLAR_E_ARG, "csr_spmm: dtype %d", dtype);
  if (n <= 0 || R <= 0 || ldx < R || ldy < R) return set_error(POLAR_E_ARG, "csr_spmm: bad n/R/ld");
  if (!indptr || !indices || !values || !X || !Y) return set_error(POLAR_E_ARG, "csr_spmm: null pointer");
  if (n > INT32_MAX) return set_error(POLAR_E_ARG, "csr_spmm: n exceeds int32 row ids");
  hipStream_t st = (hipStream_t)stream;
  // widest vector (<= 16 bytes) that divides R and keeps every row of X and Y aligned; column
  // slabs of 64 lanes cover any R
  const int esz = elem_size(dtype);
  int cpl = 16 / esz;
  while (cpl > 1 && !(R % cpl == 0 && (uintptr_t)X % (cpl * esz) == 0 && (uintptr_t)Y % (cpl * esz) == 0 &&
                      ldx % cpl == 0 && ldy % cpl == 0))
    cpl /= 2;
  if (dtype == POLAR_F64) {
    if (cpl == 2) return csr_spmm_impl<double, 2>(n, R, indptr, indices, (const double*)values, (const double*)X, ldx, (double*)Y, ldy, st);
    return csr_spmm_impl<double, 1>(n, R, indptr, indices, (const double*)values, (const double*)X, ldx, (double*)Y, ldy, st);
  }
  if (cpl == 4) return csr_spmm_impl<float, 4>(n, R, indptr, indices, (const float*)values, (const float*)X, ldx, (float*)Y, ldy, st);
  if (cpl == 2) return csr_spmm_impl<float, 2>(n, R, indptr, indices, (const float*)values, (const float*)X, ldx, (float*)Y, ldy, st);
  return csr_spmm_impl<float, 1>(n, R, indptr, indices, (const float*)values, (const float*)X, ldx, (float*)Y, ldy, st);
}

int64_t polar_msolve_ws_bytes(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags) {
  if (check_op(op)) return -1;
  return op->dtype == POLAR_F64 ? msolve_ws<double>(op, 1, R, p, nA, flags) : msolve_ws<float>(op, 1, R, p, nA, flags);
}

int polar_msolve(const polar_op* op, const void* W, int64_t ldw, int64_t R, const double* tau,
                 int32_t p, const int32_t* active, int64_t nA, double tol, int64_t maxit,
                 int32_t flags, void* Xout, int64_t x_sj, int64_t x_si, double* h_out,
                 double* te_out, int64_t* iters_out, void* ws, int64_t ws_bytes, void* stream) {
  if (int rc = check_op(op)) return rc;
  if (R <= 0 || ldw < R || p <= 0 || nA <= 0 || !W || !tau || !active || (!Xout && !(flags & POLAR_HISTORY)) || !ws)
    return set_error(POLAR_E_ARG, "bad msolve arguments");
  if (R > 512) return set_error(POLAR_E_ARG, "R=%lld exceeds 512 columns per block solve", (long long)R);
  const Segs sg = one_seg(op, nA);
  hipStream_t st = (hipStream_t)stream;
  if (op->dtype == POLAR_F64)
    return msolve_dispatch<double>(op, sg, W, ldw, R, tau, p, active, tol, maxit, flags, Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
  return msolve_dispatch<float>(op, sg, W, ldw, R, tau, p, active, tol, maxit, flags, Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
}

int64_t polar_msolve_batched_ws_bytes(const polar_op* op, const polar_segs* segs, int64_t R, int32_t p,
                                      int32_t flags) {
  if (check_op(op)) return -1;
  Segs sg;
  if (segs_from(op, segs, R, sg)) return -1;
  const int64_t nA = sg.a[sg.S];
  return op->dtype == POLAR_F64 ? msolve_ws<double>(op, sg.S, R, p, nA, flags) : msolve_ws<float>(op, sg.S, R, p, nA, flags);
}

int polar_msolve_batched(const polar_op* op, const polar_segs* segs, const void* W, int64_t ldw, int64_t R,
                         const double* tau, int32_t p, const int32_t* active, double tol, int64_t maxit,
                         int32_t flags, void* Xout, int64_t x_sj, int64_t x_si, double* h_out,
                         double* te_out, int64_t* iters_out, void* ws, int64_t ws_bytes, void* stream) {
  if (int rc = check_op(op)) return rc;
  Segs sg;
  if (int rc = segs_from(op, segs, R, sg)) return rc;
  const int64_t nA = sg.a[sg.S];
  if (R <= 0 || ldw < R || p <= 0 || !W || !tau || (nA > 0 && !active) || (!Xout && !(flags & POLAR_HISTORY)) || !ws)
    return set_error(POLAR_E_ARG, "bad msolve arguments");
  if (R > 512) return set_error(POLAR_E_ARG, "R=%lld exceeds 512 columns per block solve", (long long)R);
  if ((int64_t)sg.S * R > (int64_t)INT32_MAX / 64 || (int64_t)sg.S * R * p > (int64_t)INT32_MAX / 2)
    return set_error(POLAR_E_ARG, "too many segment columns");
  hipStream_t st = (hipStream_t)stream;
  if (op->dtype == POLAR_F64)
    return msolve_dispatch<double>(op, sg, W, ldw, R, tau, p, active, tol, maxit, flags, Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
  return msolve_dispatch<float>(op, sg, W, ldw, R, tau, p, active, tol, maxit, flags, Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
}

int64_t polar_msolve_history_ws_bytes(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags,
                                      int64_t tcap) {
  if (check_op(op) || R <= 0 || p <= 0 || nA <= 0 || tcap <= 0) return -1;
  flags |= POLAR_HISTORY;
  const bool repl = (flags & POLAR_REPLACE) != 0;
  if (op->dtype == POLAR_F64) {
    SolveMem<double> m;
    return carve_solve<double>(nullptr, op, 1, R, p, nA, window_of(flags, p), repl, m, tcap);
  }
  SolveMem<float> m;
  return carve_solve<float>(nullptr, op, 1, R, p, nA, window_of(flags, p), repl, m, tcap);
}

int polar_history_rows(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags, int64_t tcap,
                       void* ws, int64_t ws_bytes, int32_t K, const int32_t* rows, void* E, void* stream) {
  if (int rc = check_op(op)) return rc;
  if (K <= 0 || !rows || !E) return set_error(POLAR_E_ARG, "bad history_rows arguments");
  hipStream_t st = (hipStream_t)stream;
  if (op->dtype == POLAR_F64)
    return history_rows_impl<double>(op, R, p, nA, flags, tcap, ws, ws_bytes, K, rows, (double*)E, st);
  return history_rows_impl<float>(op, R, p, nA, flags, tcap, ws, ws_bytes, K, rows, (float*)E, st);
}

int polar_history_cols(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags, int64_t tcap,
                       void* ws, int64_t ws_bytes, int32_t C, const int32_t* cols, const int32_t* active,
                       void* X, void* stream) {
  if (int rc = check_op(op)) return rc;
  if (C <= 0 || !cols || !active || !X) return set_error(POLAR_E_ARG, "bad history_cols arguments");
  hipStream_t st = (hipStream_t)stream;
  if (op->dtype == POLAR_F64)
    return history_cols_impl<double>(op, R, p, nA, flags, tcap, ws, ws_bytes, C, cols, active, (double*)X, st);
  return history_cols_impl<float>(op, R, p, nA, flags, tcap, ws, ws_bytes, C, cols, active, (float*)X, st);
}

}  // extern "C"
