// Operator SpMM and the fused block multi-shift CG (gfx950, wave64).
//
// Reference semantics: realworld/spmm.py:3-25 (csr_spmm), realworld/polar_gpu.py:63-80 (Op),
// realworld/polar_gpu.py:88-133 (msolve).  Design (not a translation of the CUDA kernel):
//  * Operator stored as diagonal + off-diagonal adjacency (value-free for unit weights).  A row is
//    processed by a group of L lanes, each lane owning CPL consecutive columns (16-byte loads);
//    the group loads L column indices in one coalesced load and broadcasts them with ds_bpermute.
//  * Rows longer than POLAR_CHUNK nnz are split into chunks processed by independent groups;
//    chunk partials are combined in chunk order by a finishing kernel (deterministic).
//  * Every reduction (d.v, |r|^2) is a fixed-order two-level sum: per-workgroup partials, then the
//    last workgroup to finish sums them in block order and runs the scalar recurrences of the
//    iteration (Jegerlehner zeta, alpha/beta of every pole, freezing, convergence) in FP64.
//  * Per iteration: opmul (+dot)  [-> longfin]  -> resid (+|r|^2, scalars)  -> update (d and the
//    shifted x/p of every live pole on the active rows).  A device flag makes launches after
//    convergence no-ops, so the host reads it only every 8 iterations.
#include "common.hpp"

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
  __shared__ double red_hi[kBlock * 4];
  __shared__ double red_lo[kBlock * 4];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    red_hi[threadIdx.x * CPL + k] = acc[k].hi;
    red_lo[threadIdx.x * CPL + k] = acc[k].lo;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < R; c += kBlock) {
    int li = c / CPL, k = c % CPL;
    DD s{0.0, 0.0};
    for (int w = 0; w < kWaves; ++w)
      for (int gi = 0; gi < g.G; ++gi) {
        const int ix = ((w * 64) + gi * g.L + li) * CPL + k;
        s = dd_merge(s, DD{red_hi[ix], red_lo[ix]});
      }
    out[c] = s.hi;
    out[R + c] = s.lo;
  }
  __syncthreads();
}

// Last-workgroup election after every workgroup stored its partial row.  Returns true in the
// last workgroup (all threads), after an acquire fence.
__device__ __forceinline__ bool last_block(unsigned* counter) {
  __shared__ int s_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(counter, 1u);
    s_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (s_last) __threadfence();
  return s_last != 0;
}

// Sum the double-double partial rows parts[b*2R + (0|R) + c], b < nparts, into sh_out[c].
__device__ __forceinline__ void sum_parts(const double* __restrict__ parts, int nparts, int R,
                                          double* __restrict__ sh_out /* shared [R] */) {
  __shared__ double seg_hi[kBlock];
  __shared__ double seg_lo[kBlock];
  for (int c0 = 0; c0 < R; c0 += kBlock) {
    const int t = threadIdx.x;
    const int cw = min(R - c0, kBlock);
    const int ns = max(1, kBlock / cw);
    const int col = t % cw, seg = t / cw;
    if (seg < ns) {
      DD s{0.0, 0.0};
      for (int b = seg; b < nparts; b += ns) {
        const double* pb = parts + (int64_t)b * 2 * R;
        s = dd_merge(s, DD{pb[c0 + col], pb[R + c0 + col]});
      }
      seg_hi[seg * cw + col] = s.hi;
      seg_lo[seg * cw + col] = s.lo;
    }
    __syncthreads();
    if (t < cw) {
      DD tot{0.0, 0.0};
      for (int sgi = 0; sgi < ns; ++sgi) tot = dd_merge(tot, DD{seg_hi[sgi * cw + t], seg_lo[sgi * cw + t]});
      sh_out[c0 + t] = tot.hi + tot.lo;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------
// Gather-accumulate acc[:] += sum_k w_k X[idx_k, c0:c0+CPL] over k in [beg, end).
template <typename T, int CPL, bool WEIGHTED>
__device__ __forceinline__ void gather_acc(const int32_t* __restrict__ idx, const T* __restrict__ wt,
                                           const T* __restrict__ X, int64_t ldx, int beg, int end,
                                           const Geo& g, bool colok, int c0, T (&acc)[CPL]) {
  const int gbase = g.gi * g.L;
  for (int k0 = beg; k0 < end; k0 += g.L) {
    const int cnt = min(g.L, end - k0);
    const int myj = (g.li < cnt) ? idx[k0 + g.li] : 0;
    T myw = T(1);
    if (WEIGHTED) myw = (g.li < cnt) ? wt[k0 + g.li] : T(0);
    int t = 0;
    for (; t + 4 <= cnt; t += 4) {
      int j[4];
      T w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        j[u] = __shfl(myj, gbase + t + u, 64);
        if (WEIGHTED) w[u] = __shfl(myw, gbase + t + u, 64);
      }
      if (colok) {
        T x[4][CPL];
#pragma unroll
        for (int u = 0; u < 4; ++u) vload<T, CPL>(X + (int64_t)j[u] * ldx + c0, x[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int e = 0; e < CPL; ++e) acc[e] += (WEIGHTED ? w[u] * x[u][e] : x[u][e]);
      }
    }
    for (; t < cnt; ++t) {
      int jj = __shfl(myj, gbase + t, 64);
      T ww = T(1);
      if (WEIGHTED) ww = __shfl(myw, gbase + t, 64);
      if (colok) {
        T x[CPL];
        vload<T, CPL>(X + (int64_t)jj * ldx + c0, x);
#pragma unroll
        for (int e = 0; e < CPL; ++e) acc[e] += (WEIGHTED ? ww * x[e] : x[e]);
      }
    }
  }
}

// y = (diag + shift) x - acc + patch(x) for a complete row; returns y in yv.
template <typename T, int CPL>
__device__ __forceinline__ void finish_row(const polar_op& op, double shift, const T* __restrict__ X,
                                           int64_t ldx, int row, int c0, int pidx, const T (&acc)[CPL],
                                           T (&xv)[CPL], T (&yv)[CPL]) {
  vload<T, CPL>(X + (int64_t)row * ldx + c0, xv);
  const T dg = (T)((double)((const T*)op.diag)[row] + shift);
#pragma unroll
  for (int e = 0; e < CPL; ++e) yv[e] = (dg * xv[e]) - acc[e];
  if (pidx >= 0) {
    const int pb = op.patch_ptr[pidx], pe = op.patch_ptr[pidx + 1];
    for (int q = pb; q < pe; ++q) {
      T xj[CPL];
      vload<T, CPL>(X + (int64_t)op.patch_cols[q] * ldx + c0, xj);
#pragma unroll
      for (int e = 0; e < CPL; ++e) yv[e] = yv[e] + (xv[e] - xj[e]);
    }
  }
}

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
  unsigned* counter;   // [4] last-block counters
  unsigned long long* stats;  // [2] sum over iterations of active columns / live (pole, column) pairs
  double* parts;       // [2*kMaxGrid][2][R] double-double block partials
  int R, p;
  double tol;
  int freeze, track, trace;
};

// Phase 1 (after d.v): a = act ? rr/dv : 0.
__device__ void phase_alpha(const CGState& s, int nparts) {
  __shared__ double dv[512];
  sum_parts(s.parts, nparts, s.R, dv);
  for (int c = threadIdx.x; c < s.R; c += kBlock) s.a[c] = s.act[c] ? s.rr[c] / dv[c] : 0.0;
}

// Phase 2 (after |rn|^2): all scalar recurrences of polar_gpu.py:106-125.
__device__ void phase_scalars(const CGState& s, int nparts, int cur) {
  __shared__ double rrn[512];
  __shared__ double bec[512];
  __shared__ double ccs[512];
  __shared__ int any_s;
  sum_parts(s.parts, nparts, s.R, rrn);
  const int R = s.R;
  __shared__ int nlive_s, nact_s;
  if (threadIdx.x == 0) {
    any_s = 0;
    nlive_s = 0;
    nact_s = 0;
  }
  for (int c = threadIdx.x; c < R; c += kBlock) {
    const bool act = s.act[c] != 0;
    bec[c] = act ? rrn[c] / s.rr[c] : 0.0;
    ccs[c] = act ? s.a[c] * s.bp[c] / s.ap[c] : 0.0;
  }
  __syncthreads();
  int nlive = 0;
  for (int idx = threadIdx.x; idx < s.p * R; idx += kBlock) {
    const int c = idx % R;
    const bool act = s.act[c] != 0;
    const double a = s.a[c], be = bec[c], cc = ccs[c];
    const double z = s.z[idx], zp = s.zp[idx];
    const double zn = z / (1.0 + a * s.shift[idx / R] + cc * (1.0 - z / zp));
    const double fac = zn / z;
    const bool m = (s.live[idx] != 0) && act;
    const double aj = m ? a * fac : 0.0;
    const double bj = be * fac * fac;
    s.zn[idx] = zn;
    s.aj[idx] = aj;
    s.bj[idx] = bj;
    s.m[idx] = m ? 1 : 0;
    nlive += m ? 1 : 0;
    if (s.track) {
      const double ip = s.ip[idx], pn = s.pn[idx];
      s.h[idx] += 2.0 * aj * ip + aj * aj * pn;
      if (m) {
        s.ip[idx] = bj * (ip + aj * pn);
        s.pn[idx] = zn * zn * rrn[c] + bj * bj * pn;
      }
    }
    if (s.trace) {
      const double tg = s.tg[idx];
      s.te[idx] += aj * tg;
      if (m) s.tg[idx] = bj * tg;
    }
    if (m) {
      s.zp[idx] = z;
      s.z[idx] = zn;
    }
    const double nr = sqrt(rrn[c]);
    if (s.freeze && m && fabs(zn) * nr <= s.stop[c]) s.live[idx] = 0;
  }
  if (nlive) atomicAdd(&nlive_s, nlive);
  __syncthreads();
  int any = 0, nact = 0;
  for (int c = threadIdx.x; c < R; c += kBlock) {
    const bool act = s.act[c] != 0;
    nact += act ? 1 : 0;
    s.dmask[c] = act ? 1 : 0;
    s.be[c] = bec[c];
    if (act) {
      const double nr = sqrt(rrn[c]);
      s.iters[c] += 1;
      const bool done = nr <= s.stop[c];
      s.rr[c] = rrn[c];
      s.ap[c] = s.a[c];
      s.bp[c] = bec[c];
      if (done) s.act[c] = 0; else any = 1;
    }
  }
  if (any) atomicOr(&any_s, 1);
  if (nact) atomicAdd(&nact_s, nact);
  __syncthreads();
  if (threadIdx.x == 0) {
    s.flag[cur ^ 1] = any_s;
    s.stats[0] += (unsigned long long)nact_s;
    s.stats[1] += (unsigned long long)nlive_s;
  }
}

// ------------------------------------------------------------------------------------------------
// Kernel A: Y = Op(X) for complete rows; long-row chunk partials to lpart; optional d.v partials.
template <typename T, int CPL, bool WEIGHTED, bool DOT>
__global__ __launch_bounds__(kBlock) void k_opmul(polar_op op, double shift, const T* __restrict__ X,
                                                  int64_t ldx, T* __restrict__ Y, int64_t ldy, int R,
                                                  int L, T* __restrict__ lpart, CGState s, int cur,
                                                  int finish) {
  if (DOT && s.flag[cur] == 0) return;
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  const int64_t n = op.n;
  const int64_t nunits = n + op.n_extra;
  const int gpb = kWaves * g.G;
  int64_t u = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi;
  const int64_t stride = (int64_t)gridDim.x * gpb;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  int pc = (op.n_patch > 0 && u < n) ? lower_bound_i32(op.patch_rows, op.n_patch, (int)u) : 0;
  for (; u < nunits; u += stride) {
    int row, beg, end, lidx = -1, chunk = 0;
    if (u < n) {
      row = (int)u;
      beg = op.indptr[row];
      end = op.indptr[row + 1];
      if (end - beg > POLAR_CHUNK) {
        lidx = lower_bound_i32(op.long_rows, op.n_long, row);
        end = beg + POLAR_CHUNK;
      }
    } else {
      const int2 ex = reinterpret_cast<const int2*>(op.extra)[u - n];
      lidx = ex.x;
      chunk = ex.y;
      row = op.long_rows[lidx];
      const int rb = op.indptr[row], re = op.indptr[row + 1];
      beg = rb + chunk * POLAR_CHUNK;
      end = min(re, beg + POLAR_CHUNK);
    }
    T acc[CPL];
#pragma unroll
    for (int e = 0; e < CPL; ++e) acc[e] = T(0);
    gather_acc<T, CPL, WEIGHTED>(op.indices, (const T*)op.weights, X, ldx, beg, end, g, colok, c0, acc);
    if (lidx < 0) {
      int pidx = -1;
      if (op.n_patch > 0) {
        while (pc < op.n_patch && op.patch_rows[pc] < row) ++pc;
        if (pc < op.n_patch && op.patch_rows[pc] == row) pidx = pc;
      }
      if (colok) {
        T xv[CPL], yv[CPL];
        finish_row<T, CPL>(op, shift, X, ldx, row, c0, pidx, acc, xv, yv);
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
    if (finish && last_block(&s.counter[0])) {
      phase_alpha(s, gridDim.x);
      __syncthreads();
      if (threadIdx.x == 0) s.counter[0] = 0;
    }
  }
}

// Kernel L: combine long-row chunk partials in chunk order, finish the rows, d.v partials.
template <typename T, int CPL, bool DOT>
__global__ __launch_bounds__(kBlock) void k_longfin(polar_op op, double shift, const T* __restrict__ X,
                                                    int64_t ldx, T* __restrict__ Y, int64_t ldy, int R,
                                                    int L, const T* __restrict__ lpart, CGState s,
                                                    int cur, int parts_off) {
  if (DOT && s.flag[cur] == 0) return;
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  const int gpb = kWaves * g.G;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  for (int64_t l = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; l < op.n_long;
       l += (int64_t)gridDim.x * gpb) {
    const int row = op.long_rows[l];
    const int deg = op.indptr[row + 1] - op.indptr[row];
    const int nch = (deg + POLAR_CHUNK - 1) / POLAR_CHUNK;
    int pidx = -1;
    if (op.n_patch > 0) {
      int pc = lower_bound_i32(op.patch_rows, op.n_patch, row);
      if (pc < op.n_patch && op.patch_rows[pc] == row) pidx = pc;
    }
    if (colok) {
      T acc[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) acc[e] = T(0);
      const T* base = lpart + (int64_t)op.long_base[l] * R + c0;
      for (int ch = 0; ch < nch; ++ch) {
        T pv[CPL];
        vload<T, CPL>(base + (int64_t)ch * R, pv);
#pragma unroll
        for (int e = 0; e < CPL; ++e) acc[e] += pv[e];
      }
      T xv[CPL], yv[CPL];
      finish_row<T, CPL>(op, shift, X, ldx, row, c0, pidx, acc, xv, yv);
      vstore<T, CPL>(Y + (int64_t)row * ldy + c0, yv);
      if (DOT) {
#pragma unroll
        for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)xv[e] * (double)yv[e]);
      }
    }
  }
  if (DOT) {
    block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)(parts_off + blockIdx.x) * 2 * R);
    if (last_block(&s.counter[1])) {
      phase_alpha(s, parts_off + gridDim.x);
      __syncthreads();
      if (threadIdx.x == 0) s.counter[1] = 0;
    }
  }
}

// Kernel B: rn = r - a v (active columns), |rn|^2 partials; last block runs phase_scalars.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_resid(int64_t n, int R, int L, T* __restrict__ r,
                                                  const T* __restrict__ v, CGState s, int cur) {
  if (s.flag[cur] == 0) {
    // converged: keep both parities at 0 so every later launch is a no-op
    if (blockIdx.x == 0 && threadIdx.x == 0) s.flag[cur ^ 1] = 0;
    return;
  }
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
      act[e] = s.act[c0 + e] != 0;
      a[e] = (T)s.a[c0 + e];
      any |= act[e];
    }
    const int gpb = kWaves * g.G;
    for (int64_t i = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; i < n; i += (int64_t)gridDim.x * gpb) {
      T rv[CPL], vv[CPL];
      vload<T, CPL>(r + i * R + c0, rv);
      if (any) {
        vload<T, CPL>(v + i * R + c0, vv);
#pragma unroll
        for (int e = 0; e < CPL; ++e)
          if (act[e]) rv[e] = rv[e] - (a[e] * vv[e]);
        vstore<T, CPL>(r + i * R + c0, rv);
      }
#pragma unroll
      for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)rv[e] * (double)rv[e]);
    }
  }
  block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)blockIdx.x * 2 * R);
  if (last_block(&s.counter[2])) {
    phase_scalars(s, gridDim.x, cur);
    __syncthreads();
    if (threadIdx.x == 0) s.counter[2] = 0;
  }
}

// Kernel C: d = rn + be d (columns active at the start of the iteration) on all rows, and the
// shifted updates X_j += aj P_j; P_j = zn rn[active] + bj P_j (live poles) on the active rows.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_update(int64_t n, int64_t nA, int p, int R, int L,
                                                   const T* __restrict__ r, T* __restrict__ d,
                                                   const int32_t* __restrict__ active,
                                                   T* __restrict__ P, T* __restrict__ Xo, int64_t sj,
                                                   int64_t si, int nb_d, CGState s, int cur) {
  if (s.flag[cur] == 0) return;
  Geo g(L);
  const int c0 = g.li * CPL;
  if (c0 >= R) return;
  const int gpb = kWaves * g.G;
  if ((int)blockIdx.x < nb_d) {
    bool msk[CPL];
    T be[CPL];
    bool any = false;
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      msk[e] = s.dmask[c0 + e] != 0;
      be[e] = (T)s.be[c0 + e];
      any |= msk[e];
    }
    if (!any) return;
    for (int64_t i = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; i < n; i += (int64_t)nb_d * gpb) {
      T rv[CPL], dv[CPL];
      vload<T, CPL>(r + i * R + c0, rv);
      vload<T, CPL>(d + i * R + c0, dv);
#pragma unroll
      for (int e = 0; e < CPL; ++e)
        if (msk[e]) dv[e] = rv[e] + (be[e] * dv[e]);
      vstore<T, CPL>(d + i * R + c0, dv);
    }
    return;
  }
  const int nb_s = gridDim.x - nb_d;
  for (int64_t i = (int64_t)(blockIdx.x - nb_d) * gpb + g.wave * g.G + g.gi; i < nA;
       i += (int64_t)nb_s * gpb) {
    T rn[CPL];
    vload<T, CPL>(r + (int64_t)active[i] * R + c0, rn);
    for (int j = 0; j < p; ++j) {
      const int base = j * R + c0;
      bool mj[CPL];
      bool anym = false;
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        mj[e] = s.m[base + e] != 0;
        anym |= mj[e];
      }
      if (!anym) continue;
      T aj[CPL], bj[CPL], zn[CPL];
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        aj[e] = (T)s.aj[base + e];
        bj[e] = (T)s.bj[base + e];
        zn[e] = (T)s.zn[base + e];
      }
      T* pp = P + ((int64_t)j * nA + i) * R + c0;
      T* xp = Xo + (int64_t)j * sj + i * si + c0;
      T pv[CPL], xv[CPL];
      vload<T, CPL>(pp, pv);
      vload<T, CPL>(xp, xv);
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        if (mj[e]) {
          xv[e] = xv[e] + (aj[e] * pv[e]);
          pv[e] = (zn[e] * rn[e]) + (bj[e] * pv[e]);
        }
      }
      vstore<T, CPL>(xp, xv);
      vstore<T, CPL>(pp, pv);
    }
  }
}

// Init: r = d = W[:, :R]; |w|^2 partials; last block initialises every scalar.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_init(int64_t n, int R, int L, const T* __restrict__ W,
                                                 int64_t ldw, T* __restrict__ r, T* __restrict__ d,
                                                 CGState s) {
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool colok = c0 < R;
  DD dacc[CPL];
#pragma unroll
  for (int e = 0; e < CPL; ++e) dacc[e] = DD{0.0, 0.0};
  if (colok) {
    const int gpb = kWaves * g.G;
    for (int64_t i = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; i < n; i += (int64_t)gridDim.x * gpb) {
      T wv[CPL];
      vload<T, CPL>(W + i * ldw + c0, wv);
      vstore<T, CPL>(r + i * R + c0, wv);
      vstore<T, CPL>(d + i * R + c0, wv);
#pragma unroll
      for (int e = 0; e < CPL; ++e) dd_add(dacc[e], (double)wv[e] * (double)wv[e]);
    }
  }
  block_colsum<CPL>(g, dacc, R, s.parts + (int64_t)blockIdx.x * 2 * R);
  if (last_block(&s.counter[3])) {
    __shared__ double rr[512];
    __shared__ int any_s;
    sum_parts(s.parts, gridDim.x, R, rr);
    if (threadIdx.x == 0) any_s = 0;
    __syncthreads();
    int any = 0;
    for (int c = threadIdx.x; c < R; c += kBlock) {
      s.rr[c] = rr[c];
      s.stop[c] = s.tol * sqrt(rr[c]);
      s.act[c] = rr[c] > 0.0 ? 1 : 0;
      s.dmask[c] = 0;
      s.a[c] = 0.0;
      s.be[c] = 0.0;
      s.ap[c] = 1.0;
      s.bp[c] = 0.0;
      s.iters[c] = 0;
      any |= (rr[c] > 0.0);
    }
    for (int idx = threadIdx.x; idx < s.p * R; idx += kBlock) {
      const int c = idx % R;
      s.z[idx] = 1.0;
      s.zp[idx] = 1.0;
      s.zn[idx] = 1.0;
      s.aj[idx] = 0.0;
      s.bj[idx] = 0.0;
      s.live[idx] = rr[c] > 0.0 ? 1 : 0;
      s.m[idx] = 0;
      s.h[idx] = 0.0;
      s.ip[idx] = 0.0;
      s.pn[idx] = rr[c];
      s.te[idx] = 0.0;
      s.tg[idx] = rr[c];
    }
    if (any) atomicOr(&any_s, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
      s.flag[0] = any_s;
      s.flag[1] = any_s;
      s.counter[3] = 0;
    }
  }
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
    for (int j = 0; j < p; ++j) {
      vstore<T, CPL>(P + ((int64_t)j * nA + i) * R + c0, wv);
      vstore<T, CPL>(Xo + (int64_t)j * sj + i * si + c0, zero);
    }
  }
}

// shift[j] = tau[j] - tau[0]
__global__ void k_shift(const double* __restrict__ tau, int p, double* __restrict__ shift) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < p) shift[j] = tau[j] - tau[0];
}

// ------------------------------------------------------------------------------------------------
// Host side
// Live kernel profiling (HIP events around each launch of the CG iteration), off by default.
// Class 0: opmul (+ long-row finish), 1: resid, 2: update; class 3 counts solves and, in `bytes`,
// the sum over iterations of live (pole, column) pairs.
struct ProfClass {
  double ms = 0.0;
  int64_t launches = 0;
  double bytes = 0.0;
};
static ProfClass g_prof[4];
static int g_prof_on = 0;

struct Plan {
  int cpl, L, G;
  int grid_a, grid_l, grid_rows;
};

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

static Plan make_plan(const polar_op* op, int64_t R, int cpl) {
  Plan pl;
  pl.cpl = cpl;
  pl.L = lanes_for(R, cpl);
  pl.G = 64 / pl.L;
  const int64_t gpb = (int64_t)kWaves * pl.G;
  const int64_t units = op->n + op->n_extra;
  pl.grid_a = (int)std::min<int64_t>(kMaxGrid, std::max<int64_t>(1, (units + gpb - 1) / gpb));
  pl.grid_l = (int)std::min<int64_t>(kMaxGrid, std::max<int64_t>(1, ((int64_t)op->n_long + gpb - 1) / gpb));
  pl.grid_rows = (int)std::min<int64_t>(kMaxGrid, std::max<int64_t>(1, (op->n + gpb - 1) / gpb));
  return pl;
}

static int pick_cpl(int dtype, int64_t R, std::initializer_list<std::pair<const void*, int64_t>> ptrs) {
  const int esz = elem_size(dtype);
  const int full = 16 / esz;
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

template <typename T>
static void carve_state(Carver& cv, CGState& s, int64_t R, int p) {
  s.rr = cv.take<double>(R); s.stop = cv.take<double>(R); s.a = cv.take<double>(R);
  s.be = cv.take<double>(R); s.ap = cv.take<double>(R); s.bp = cv.take<double>(R);
  s.act = cv.take<int>(R); s.dmask = cv.take<int>(R); s.iters = cv.take<int64_t>(R);
  const int64_t pr = (int64_t)p * R;
  s.z = cv.take<double>(pr); s.zp = cv.take<double>(pr); s.zn = cv.take<double>(pr);
  s.aj = cv.take<double>(pr); s.bj = cv.take<double>(pr); s.h = cv.take<double>(pr);
  s.ip = cv.take<double>(pr); s.pn = cv.take<double>(pr); s.te = cv.take<double>(pr);
  s.tg = cv.take<double>(pr); s.shift = cv.take<double>(p);
  s.live = cv.take<int>(pr); s.m = cv.take<int>(pr);
  s.flag = cv.take<int>(2);
  s.counter = cv.take<unsigned>(4);
  s.stats = cv.take<unsigned long long>(2);
  s.parts = cv.take<double>((int64_t)4 * kMaxGrid * R);
}

template <typename T, int CPL>
static int launch_opmul(const polar_op* op, const Plan& pl, double shift, const T* X, int64_t ldx, T* Y,
                        int64_t ldy, int R, T* lpart, const CGState& s, int cur, bool dot, hipStream_t st) {
  const int finishA = (op->n_long == 0) ? 1 : 0;
  if (dot) {
    if (op->weighted)
      hipLaunchKernelGGL((k_opmul<T, CPL, true, true>), dim3(pl.grid_a), dim3(kBlock), 0, st, *op, shift, X, ldx, Y, ldy, R, pl.L, lpart, s, cur, finishA);
    else
      hipLaunchKernelGGL((k_opmul<T, CPL, false, true>), dim3(pl.grid_a), dim3(kBlock), 0, st, *op, shift, X, ldx, Y, ldy, R, pl.L, lpart, s, cur, finishA);
  } else {
    if (op->weighted)
      hipLaunchKernelGGL((k_opmul<T, CPL, true, false>), dim3(pl.grid_a), dim3(kBlock), 0, st, *op, shift, X, ldx, Y, ldy, R, pl.L, lpart, s, cur, 0);
    else
      hipLaunchKernelGGL((k_opmul<T, CPL, false, false>), dim3(pl.grid_a), dim3(kBlock), 0, st, *op, shift, X, ldx, Y, ldy, R, pl.L, lpart, s, cur, 0);
  }
  POLAR_LAUNCH_CHECK();
  if (op->n_long > 0) {
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
  return launch_opmul<T, CPL>(op, pl, shift, X, ldx, Y, ldy, (int)R, lpart, s, 0, false, st);
}

// ---------------------------------------------------------------- msolve
template <typename T, int CPL>
static int msolve_impl(const polar_op* op, const T* W, int64_t ldw, int64_t R, const double* tau, int p,
                       const int32_t* active, int64_t nA, double tol, int64_t maxit, int flags, T* Xo,
                       int64_t sj, int64_t si, double* h_out, double* te_out, int64_t* iters_out,
                       void* ws, int64_t ws_bytes, hipStream_t st) {
  Plan pl = make_plan(op, R, CPL);
  const int64_t n = op->n;
  Carver cv(ws);
  CGState s{};
  carve_state<T>(cv, s, R, p);
  T* r = cv.take<T>(n * R);
  T* d = cv.take<T>(n * R);
  T* v = cv.take<T>(n * R);
  T* P = cv.take<T>((int64_t)p * nA * R);
  T* lpart = cv.take<T>((int64_t)op->n_partial * R);
  if (cv.off > ws_bytes) return set_error(POLAR_E_ARG, "msolve workspace too small (%lld < %lld)", (long long)ws_bytes, (long long)cv.off);
  s.R = (int)R;
  s.p = p;
  s.tol = tol;
  s.freeze = (flags & POLAR_NOFREEZE) ? 0 : 1;
  s.track = (flags & POLAR_TRACK) ? 1 : 0;
  s.trace = (flags & POLAR_TRACE) ? 1 : 0;
  const int L = pl.L;
  const int64_t gpb = (int64_t)kWaves * pl.G;

  POLAR_HIP(hipMemsetAsync(s.counter, 0, 4 * sizeof(unsigned), st));
  POLAR_HIP(hipMemsetAsync(s.stats, 0, 2 * sizeof(unsigned long long), st));
  const bool prof = g_prof_on != 0;
  std::vector<hipEvent_t> ev;
  std::vector<float> kms;  // per iteration: opmul, resid, update milliseconds
  if (prof) ev.reserve(4 * 8);
  hipLaunchKernelGGL(k_shift, dim3((p + 255) / 256), dim3(256), 0, st, tau, p, s.shift);
  POLAR_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_init<T, CPL>), dim3(pl.grid_rows), dim3(kBlock), 0, st, n, (int)R, L, W, ldw, r, d, s);
  POLAR_LAUNCH_CHECK();
  const int grid_sh = (int)std::min<int64_t>(65536, std::max<int64_t>(1, (nA + gpb - 1) / gpb));
  hipLaunchKernelGGL((k_init_shift<T, CPL>), dim3(grid_sh), dim3(kBlock), 0, st, nA, p, (int)R, L, W, ldw, active, P, Xo, sj, si);
  POLAR_LAUNCH_CHECK();

  const double shift0 = 0.0;  // tau[0] is applied through diag + shift below
  double tau0 = 0.0;
  POLAR_HIP(hipMemcpyAsync(&tau0, tau, sizeof(double), hipMemcpyDeviceToHost, st));
  POLAR_HIP(hipStreamSynchronize(st));
  (void)shift0;
  const int nb_d = (int)std::min<int64_t>(65536, std::max<int64_t>(1, (n + gpb - 1) / gpb));
  const int nb_s = grid_sh;
  bool converged = false;
  int64_t it = 0;
  for (; it < maxit; ++it) {
    const int cur = (int)(it & 1);
    const int slot = (int)(it % 8) * 4;
    if (prof) {
      while ((int)ev.size() < slot + 4) {
        hipEvent_t e;
        POLAR_HIP(hipEventCreate(&e));
        ev.push_back(e);
      }
      POLAR_HIP(hipEventRecord(ev[slot], st));
    }
    int rc = launch_opmul<T, CPL>(op, pl, tau0, d, R, v, R, (int)R, lpart, s, cur, true, st);
    if (rc) return rc;
    if (prof) POLAR_HIP(hipEventRecord(ev[slot + 1], st));
    hipLaunchKernelGGL((k_resid<T, CPL>), dim3(pl.grid_rows), dim3(kBlock), 0, st, n, (int)R, L, r, (const T*)v, s, cur);
    POLAR_LAUNCH_CHECK();
    if (prof) POLAR_HIP(hipEventRecord(ev[slot + 2], st));
    hipLaunchKernelGGL((k_update<T, CPL>), dim3(nb_d + nb_s), dim3(kBlock), 0, st, n, nA, p, (int)R, L,
                       (const T*)r, d, active, P, Xo, sj, si, nb_d, s, cur);
    POLAR_LAUNCH_CHECK();
    if (prof) POLAR_HIP(hipEventRecord(ev[slot + 3], st));
    if (it % 8 == 7) {
      int flag = 1;
      POLAR_HIP(hipMemcpyAsync(&flag, s.flag + ((it + 1) & 1), sizeof(int), hipMemcpyDeviceToHost, st));
      POLAR_HIP(hipStreamSynchronize(st));
      if (prof) {
        for (int k = 0; k < 8; ++k) {
          for (int c = 0; c < 3; ++c) {
            float ms = 0.f;
            POLAR_HIP(hipEventElapsedTime(&ms, ev[k * 4 + c], ev[k * 4 + c + 1]));
            kms.push_back(ms);
          }
        }
      }
      if (!flag) {
        converged = true;
        break;
      }
    }
  }
  if (prof) {
    // count only the iterations that did work (launches after convergence return at once)
    std::vector<int64_t> ih(R);
    unsigned long long st2[2];
    POLAR_HIP(hipMemcpyAsync(ih.data(), s.iters, sizeof(int64_t) * R, hipMemcpyDeviceToHost, st));
    POLAR_HIP(hipMemcpyAsync(st2, s.stats, sizeof(st2), hipMemcpyDeviceToHost, st));
    POLAR_HIP(hipStreamSynchronize(st));
    int64_t real = 0;
    for (int64_t c = 0; c < R; ++c) real = std::max(real, ih[c]);
    real = std::min<int64_t>(real, (int64_t)kms.size() / 3);
    const double esz = sizeof(T);
    const double nnz = (double)((int64_t)0);
    (void)nnz;
    double opbytes = 0.0;
    {
      int32_t nnz_h = 0;
      POLAR_HIP(hipMemcpy(&nnz_h, op->indptr + op->n, sizeof(int32_t), hipMemcpyDeviceToHost));
      opbytes = 4.0 * (n + 1) + 4.0 * nnz_h + (op->weighted ? esz * nnz_h : 0.0) + esz * n + 2.0 * esz * n * R;
    }
    for (int64_t k = 0; k < real; ++k) {
      g_prof[0].ms += kms[3 * k];
      g_prof[1].ms += kms[3 * k + 1];
      g_prof[2].ms += kms[3 * k + 2];
    }
    g_prof[0].launches += real;
    g_prof[1].launches += real;
    g_prof[2].launches += real;
    g_prof[0].bytes += opbytes * real;
    g_prof[1].bytes += 3.0 * esz * n * (double)st2[0];
    g_prof[2].bytes += esz * (3.0 * n + nA) * (double)st2[0] + 4.0 * esz * nA * (double)st2[1];
    g_prof[3].launches += 1;
    g_prof[3].bytes += (double)st2[1];
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  if (h_out) POLAR_HIP(hipMemcpyAsync(h_out, s.h, sizeof(double) * p * R, hipMemcpyDeviceToDevice, st));
  if (te_out) POLAR_HIP(hipMemcpyAsync(te_out, s.te, sizeof(double) * p * R, hipMemcpyDeviceToDevice, st));
  if (iters_out) POLAR_HIP(hipMemcpyAsync(iters_out, s.iters, sizeof(int64_t) * R, hipMemcpyDeviceToDevice, st));
  if (!converged) return set_error(POLAR_E_ITERCAP, "multi-shift CG iteration cap");
  return 0;
}

template <typename T>
static int64_t msolve_ws(const polar_op* op, int64_t R, int p, int64_t nA) {
  Carver cv(nullptr);
  CGState s{};
  carve_state<T>(cv, s, R, p);
  cv.take<T>(op->n * R);
  cv.take<T>(op->n * R);
  cv.take<T>(op->n * R);
  cv.take<T>((int64_t)p * nA * R);
  cv.take<T>((int64_t)op->n_partial * R);
  return cv.off + 256;
}

}  // namespace polar

using namespace polar;

extern "C" {

int polar_abi_version(void) { return POLAR_ABI_VERSION; }

int polar_prof_set(int32_t on) {
  g_prof_on = on;
  for (auto& c : g_prof) c = ProfClass();
  return 0;
}

int polar_prof_get(int32_t cls, double* ms, int64_t* launches, double* bytes) {
  if (cls < 0 || cls > 3) return set_error(POLAR_E_ARG, "bad profile class");
  if (ms) *ms = g_prof[cls].ms;
  if (launches) *launches = g_prof[cls].launches;
  if (bytes) *bytes = g_prof[cls].bytes;
  return 0;
}

const char* polar_last_error(void) { return g_err; }

int64_t polar_spmm_ws_bytes(const polar_op* op, int64_t R) {
  if (check_op(op)) return -1;
  return align_up((int64_t)op->n_partial * R * elem_size(op->dtype), 256) + 512;
}

int polar_spmm(const polar_op* op, double shift, const void* X, int64_t ldx, void* Y, int64_t ldy,
               int64_t R, void* ws, int64_t ws_bytes, void* stream) {
  if (int rc = check_op(op)) return rc;
  if (R <= 0 || ldx < R || ldy < R) return set_error(POLAR_E_ARG, "bad R/ld");
  hipStream_t st = (hipStream_t)stream;
  const int cpl = pick_cpl(op->dtype, R, {{X, ldx}, {Y, ldy}, {ws, 0}});
  if (lanes_for(R, cpl) > 64) return set_error(POLAR_E_ARG, "R=%lld too wide for one wave (max %d)", (long long)R, 64 * cpl);
  if (op->dtype == POLAR_F64) {
    if (cpl == 2) return spmm_impl<double, 2>(op, shift, (const double*)X, ldx, (double*)Y, ldy, R, ws, ws_bytes, st);
    return spmm_impl<double, 1>(op, shift, (const double*)X, ldx, (double*)Y, ldy, R, ws, ws_bytes, st);
  }
  if (cpl == 4) return spmm_impl<float, 4>(op, shift, (const float*)X, ldx, (float*)Y, ldy, R, ws, ws_bytes, st);
  if (cpl == 2) return spmm_impl<float, 2>(op, shift, (const float*)X, ldx, (float*)Y, ldy, R, ws, ws_bytes, st);
  return spmm_impl<float, 1>(op, shift, (const float*)X, ldx, (float*)Y, ldy, R, ws, ws_bytes, st);
}

int64_t polar_msolve_ws_bytes(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags) {
  (void)flags;
  if (check_op(op)) return -1;
  return op->dtype == POLAR_F64 ? msolve_ws<double>(op, R, p, nA) : msolve_ws<float>(op, R, p, nA);
}

int polar_msolve(const polar_op* op, const void* W, int64_t ldw, int64_t R, const double* tau,
                 int32_t p, const int32_t* active, int64_t nA, double tol, int64_t maxit,
                 int32_t flags, void* Xout, int64_t x_sj, int64_t x_si, double* h_out,
                 double* te_out, int64_t* iters_out, void* ws, int64_t ws_bytes, void* stream) {
  if (int rc = check_op(op)) return rc;
  if (R <= 0 || ldw < R || p <= 0 || nA <= 0 || !W || !tau || !active || !Xout || !ws)
    return set_error(POLAR_E_ARG, "bad msolve arguments");
  if (R > 512) return set_error(POLAR_E_ARG, "R=%lld exceeds 512 columns per block solve", (long long)R);
  hipStream_t st = (hipStream_t)stream;
  const int cpl = pick_cpl(op->dtype, R, {{W, ldw}, {Xout, x_si}, {(const char*)Xout + 0, x_sj}});
  if (lanes_for(R, cpl) > 64) return set_error(POLAR_E_ARG, "R=%lld too wide for one wave at this alignment", (long long)R);
  if (op->dtype == POLAR_F64) {
    if (cpl == 2) return msolve_impl<double, 2>(op, (const double*)W, ldw, R, tau, p, active, nA, tol, maxit, flags, (double*)Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
    return msolve_impl<double, 1>(op, (const double*)W, ldw, R, tau, p, active, nA, tol, maxit, flags, (double*)Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
  }
  if (cpl == 4) return msolve_impl<float, 4>(op, (const float*)W, ldw, R, tau, p, active, nA, tol, maxit, flags, (float*)Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
  if (cpl == 2) return msolve_impl<float, 2>(op, (const float*)W, ldw, R, tau, p, active, nA, tol, maxit, flags, (float*)Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
  return msolve_impl<float, 1>(op, (const float*)W, ldw, R, tau, p, active, nA, tol, maxit, flags, (float*)Xout, x_sj, x_si, h_out, te_out, iters_out, ws, ws_bytes, st);
}

}  // extern "C"
