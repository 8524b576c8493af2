// Greedy-round kernels: endpoint-sketch bank gathers, proxy scores, top-B shortlist, re-rank,
// Sherman-Morrison cache maintenance and the counter-hash Rademacher sketches (gfx950, wave64).
//
// Reference semantics: realworld/polar_gpu.py:200-251 (bank, proxy, shortlist, re-rank,
// maintenance) and :148-165 (sketches).  All of these are HBM-bound streams or gathers over
// (p x N) or (p x nA x q) arrays; rows of q sketch columns are read by a group of lanes with
// 16-byte loads and reduced across the group with ds_swizzle/DPP (__shfl_xor).
#include "common.hpp"

namespace polar {

// Sum over the L lanes of a row-group (L power of two); result valid in every lane of the group.
__device__ __forceinline__ double group_sum(double v, int L) {
  for (int o = L >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T, int CPL>
__device__ __forceinline__ void load_cols(const T* __restrict__ row, int c0, int64_t q, double (&o)[CPL]) {
  if (c0 + CPL <= q) {
    T t[CPL];
    vload<T, CPL>(row + c0, t);
#pragma unroll
    for (int e = 0; e < CPL; ++e) o[e] = (double)t[e];
  } else {
#pragma unroll
    for (int e = 0; e < CPL; ++e) o[e] = (c0 + e < q) ? (double)row[c0 + e] : 0.0;
  }
}

// ------------------------------------------------------------------------------------------------
// out[j*N + e] = |X[j, ea[e], :] - X[j, eb[e], :]|^2
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_pair_sqnorm(const T* __restrict__ X, int64_t sj, int64_t si,
                                                        int p, int64_t q, const int32_t* __restrict__ ea,
                                                        const int32_t* __restrict__ eb, int64_t N,
                                                        double* __restrict__ out, int L) {
  Geo g(L);
  const int c0 = g.li * CPL;
  const int64_t gpb = (int64_t)kWaves * g.G;
  const int64_t total = (int64_t)p * N;
  for (int64_t item = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; item < total;
       item += (int64_t)gridDim.x * gpb) {
    const int j = (int)(item / N);
    const int64_t e = item - (int64_t)j * N;
    const T* xa = X + (int64_t)j * sj + (int64_t)ea[e] * si;
    const T* xb = X + (int64_t)j * sj + (int64_t)eb[e] * si;
    DD s{0.0, 0.0};
    if (c0 < q) {
      double va[CPL], vb[CPL];
      load_cols<T, CPL>(xa, c0, q, va);
      load_cols<T, CPL>(xb, c0, q, vb);
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const double df = va[k] - vb[k];
        dd_add(s, df * df);
      }
    }
    s = dd_group_sum(s, L);
    if (g.li == 0) out[item] = s.hi + s.lo;
  }
}

// out[e] (+)= sum_j om_j |X[j, ea[e], :] - X[j, eb[e], :]|^2 (pole-weighted pair norms; the
// adjoint of the sketched trace with respect to the weight of pair e is minus this sum).
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_pair_wsum(const T* __restrict__ X, int64_t sj, int64_t si, int p,
                                                      int64_t q, const int32_t* __restrict__ ea,
                                                      const int32_t* __restrict__ eb, int64_t E,
                                                      const double* __restrict__ om, double* __restrict__ out,
                                                      int accumulate, int L) {
  Geo g(L);
  const int c0 = g.li * CPL;
  const int64_t gpb = (int64_t)kWaves * g.G;
  for (int64_t e = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi; e < E; e += (int64_t)gridDim.x * gpb) {
    const int64_t a = ea[e], b = eb[e];
    DD tot{0.0, 0.0};
    for (int j = 0; j < p; ++j) {
      DD s{0.0, 0.0};
      if (c0 < q) {
        double va[CPL], vb[CPL];
        load_cols<T, CPL>(X + (int64_t)j * sj + a * si, c0, q, va);
        load_cols<T, CPL>(X + (int64_t)j * sj + b * si, c0, q, vb);
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          const double df = va[k] - vb[k];
          dd_add(s, df * df);
        }
      }
      s = dd_group_sum(s, L);
      dd_add(tot, om[j] * (s.hi + s.lo));
    }
    if (g.li == 0) out[e] = (accumulate ? out[e] : 0.0) + (tot.hi + tot.lo);
  }
}

// score[e] = sum_j om_j h / (1 + t + (sigma + tau_j) h)  | log1p(t0 + sigma h0); dead -> -inf
__global__ __launch_bounds__(kBlock) void k_proxy(int obj, const double* __restrict__ h,
                                                  const double* __restrict__ t, const double* __restrict__ tau,
                                                  const double* __restrict__ om, int p, int64_t N, double sigma,
                                                  const uint8_t* __restrict__ dead, double* __restrict__ score) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (int64_t)gridDim.x * blockDim.x) {
    double sc;
    if (obj == 2) {
      sc = log1p(t[e] + sigma * h[e]);
    } else if (obj == 3) {
      // first-order gain: minus the derivative of the sketched trace at weight 0
      sc = 0.0;
      for (int j = 0; j < p; ++j) sc = sc + om[j] * h[(int64_t)j * N + e];
    } else {
      sc = 0.0;
      for (int j = 0; j < p; ++j) {
        const double hv = h[(int64_t)j * N + e], tv = t[(int64_t)j * N + e];
        sc = sc + ((om[j] * hv) / ((1.0 + tv) + ((sigma + tau[j]) * hv)));
      }
    }
    if (dead && dead[e]) sc = -INFINITY;
    score[e] = sc;
  }
}

// ------------------------------------------------------------------------------------------------
// Top-B in the order (score desc, index asc): pass b selects the maximum among elements that come
// strictly after the pass b-1 pick.  sel_s/sel_i: device slots of previous picks.
__device__ __forceinline__ bool better(double s1, int64_t i1, double s2, int64_t i2) {
  // true if (s1, i1) precedes (s2, i2) in (score desc, index asc)
  return (s1 > s2) || (s1 == s2 && i1 < i2);
}

__device__ __forceinline__ void best_tree(double* ss, int64_t* si) {
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const int64_t i2 = si[threadIdx.x + o];
      if (i2 >= 0 && (si[threadIdx.x] < 0 || better(ss[threadIdx.x + o], i2, ss[threadIdx.x], si[threadIdx.x]))) {
        ss[threadIdx.x] = ss[threadIdx.x + o];
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
}

// Pass `pass`, stage 1: best element (after the previous pick) of each workgroup's slice.
__global__ __launch_bounds__(kBlock) void k_select_blocks(const double* __restrict__ score, int64_t N, int pass,
                                                          const double* __restrict__ sel_s,
                                                          const int32_t* __restrict__ out, double* __restrict__ bs,
                                                          int64_t* __restrict__ bi) {
  const bool has_prev = pass > 0;
  const double ps = has_prev ? sel_s[pass - 1] : 0.0;
  const int64_t pi = has_prev ? (int64_t)out[pass - 1] : -1;
  double best_s = -INFINITY;
  int64_t best_i = -1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (int64_t)gridDim.x * blockDim.x) {
    const double sc = score[e];
    const bool after = !has_prev || better(ps, pi, sc, e);
    if (after && (best_i < 0 || better(sc, e, best_s, best_i))) {
      best_s = sc;
      best_i = e;
    }
  }
  __shared__ double ss[kBlock];
  __shared__ int64_t si[kBlock];
  ss[threadIdx.x] = best_s;
  si[threadIdx.x] = best_i;
  best_tree(ss, si);
  if (threadIdx.x == 0) {
    bs[blockIdx.x] = ss[0];
    bi[blockIdx.x] = si[0];
  }
}

// Stage 2 (one workgroup): best of the workgroup results -> out[pass], sel_s[pass].
__global__ __launch_bounds__(kBlock) void k_select_final(int pass, int nb, const double* __restrict__ bs,
                                                         const int64_t* __restrict__ bi, double* __restrict__ sel_s,
                                                         int32_t* __restrict__ out) {
  __shared__ double ss[kBlock];
  __shared__ int64_t si[kBlock];
  ss[threadIdx.x] = (int)threadIdx.x < nb ? bs[threadIdx.x] : -INFINITY;
  si[threadIdx.x] = (int)threadIdx.x < nb ? bi[threadIdx.x] : -1;
  best_tree(ss, si);
  if (threadIdx.x == 0) {
    sel_s[pass] = ss[0];
    out[pass] = (int32_t)si[0];
  }
}

// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void k_pair_rhs(int Br, const int32_t* __restrict__ a, const int32_t* __restrict__ b, T* __restrict__ W) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < Br) {
    W[(int64_t)a[c] * Br + c] = T(1);
    W[(int64_t)b[c] * Br + c] = T(-1);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_rerank(int obj, const T* __restrict__ U, int64_t u_sj, int64_t u_si,
                                                   int p, int Br, const int32_t* __restrict__ sa,
                                                   const int32_t* __restrict__ sb, const double* __restrict__ hh,
                                                   const double* __restrict__ om, double* __restrict__ g,
                                                   double* __restrict__ res) {
  for (int c = threadIdx.x; c < Br; c += blockDim.x) {
    double acc = 0.0;
    if (obj == 2) {
      const double bu = (double)U[(int64_t)sa[c] * u_si + c] - (double)U[(int64_t)sb[c] * u_si + c];
      acc = log1p(bu);
    } else {
      for (int j = 0; j < p; ++j) {
        const T* Uj = U + (int64_t)j * u_sj;
        const double bu = (double)Uj[(int64_t)sa[c] * u_si + c] - (double)Uj[(int64_t)sb[c] * u_si + c];
        acc = acc + ((om[j] * hh[(int64_t)j * Br + c]) / (1.0 + bu));
      }
    }
    g[c] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    for (int c = 1; c < Br; ++c)
      if (g[c] > g[best]) best = c;
    res[0] = (double)best;
    res[1] = g[best];
  }
}

// ------------------------------------------------------------------------------------------------
// Maintenance.  coef per pole j: [xi(q) | eta(q) | D | nxi | neta]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_maint_prep(const T* __restrict__ X, const T* __restrict__ Y, int64_t sj,
                                                       int64_t si, int64_t q, const T* __restrict__ u, int64_t u_sj,
                                                       int64_t u_si, int aa, int bb, const double* __restrict__ gc,
                                                       double* __restrict__ coef) {
  const int j = blockIdx.x;
  const int64_t stride = 2 * q + 3;
  double* cj = coef + (int64_t)j * stride;
  const T* Xa = X + (int64_t)j * sj + (int64_t)aa * si;
  const T* Xb = X + (int64_t)j * sj + (int64_t)bb * si;
  const T* Ya = Y + (int64_t)j * sj + (int64_t)aa * si;
  const T* Yb = Y + (int64_t)j * sj + (int64_t)bb * si;
  __shared__ DD sx[kBlock], se[kBlock];
  DD ax{0.0, 0.0}, ae{0.0, 0.0};
  for (int64_t c = threadIdx.x; c < q; c += blockDim.x) {
    const double xi = (double)Xa[c] - (double)Xb[c];
    const double eta = gc[c] - ((double)Ya[c] - (double)Yb[c]);
    cj[c] = xi;
    cj[q + c] = eta;
    dd_add(ax, xi * xi);
    dd_add(ae, eta * eta);
  }
  sx[threadIdx.x] = ax;
  se[threadIdx.x] = ae;
  __syncthreads();
  if (threadIdx.x == 0) {
    DD tx{0.0, 0.0}, te{0.0, 0.0};
    for (int k = 0; k < (int)blockDim.x; ++k) {
      tx = dd_merge(tx, sx[k]);
      te = dd_merge(te, se[k]);
    }
    const T* uj = u + (int64_t)j * u_sj;
    cj[2 * q] = (1.0 + (double)uj[(int64_t)aa * u_si]) - (double)uj[(int64_t)bb * u_si];
    cj[2 * q + 1] = tx.hi + tx.lo;
    cj[2 * q + 2] = te.hi + te.lo;
  }
}

// One pass over the old rows: w = row . coef, then row -/+= u * coef / D (X with xi, Y with eta).
// FP64 rows apply u * coef / D exactly as written (the pinned arithmetic); FP32 rows multiply by
// u / D computed once per item (the double-precision difference is far below FP32 rounding).
// Each row of X and Y is read once and written once per call, so both go through non-temporal
// loads and stores (kept out of L2).  A row group loads its next item before it reduces and
// updates the current one, so a wave keeps two rows of X and Y in flight.
template <typename T, int CPL>
struct RankRow {
  T xv[CPL], yv[CPL];
  double ui;
};

template <typename T, int CPL>
__device__ __forceinline__ void rank_load(const T* __restrict__ X, const T* __restrict__ Y, int64_t sj, int64_t si,
                                          int64_t q, const T* __restrict__ u, int64_t u_sj, int64_t u_si, int j,
                                          int64_t i, int c0, RankRow<T, CPL>& r) {
  r.ui = (double)u[(int64_t)j * u_sj + i * u_si];
  const T* xr = X + (int64_t)j * sj + i * si;
  const T* yr = Y + (int64_t)j * sj + i * si;
  if (c0 + CPL <= q) {
    vload_nt<T, CPL>(xr + c0, r.xv);
    vload_nt<T, CPL>(yr + c0, r.yv);
  } else {
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      r.xv[e] = (c0 + e < q) ? xr[c0 + e] : T(0);
      r.yv[e] = (c0 + e < q) ? yr[c0 + e] : T(0);
    }
  }
}

// The lane's xi/eta entries and D of one pole, reloaded only when the pole changes.
template <int CPL>
struct PoleCoef {
  int j = -1;
  double cx[CPL], cy[CPL], D;
  __device__ __forceinline__ void load(const double* __restrict__ coef, int64_t q, int jn, int c0) {
    j = jn;
    const double* cj = coef + (int64_t)jn * (2 * q + 3);
    D = cj[2 * q];
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      cx[e] = (c0 + e < q) ? cj[c0 + e] : 0.0;
      cy[e] = (c0 + e < q) ? cj[q + c0 + e] : 0.0;
    }
  }
};

// Reduces and updates the item (pole j, active row i) held in `cur`.  Every lane of the row group
// takes part (the dots are butterflies over its L lanes); `live` is false for the lanes past the
// last column and for a row group past the last row, which then store nothing.
template <typename T, int CPL>
__device__ __forceinline__ void rank_item(T* __restrict__ X, T* __restrict__ Y, int64_t sj, int64_t si, int64_t q,
                                          int L, int c0, bool live, int j, int64_t i, const RankRow<T, CPL>& cur,
                                          const PoleCoef<CPL>& pc, double ui, double& wx, double& wy) {
  if constexpr (sizeof(T) == 8) {  // FP64 rows: double-double dots, rounded once (pinned)
    DD sx{0.0, 0.0}, sy{0.0, 0.0};
    if (live) {
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        if (c0 + e < q) {
          dd_add(sx, (double)cur.xv[e] * pc.cx[e]);
          dd_add(sy, (double)cur.yv[e] * pc.cy[e]);
        }
      }
    }
    sx = dd_group_sum(sx, L);
    sy = dd_group_sum(sy, L);
    wx = sx.hi + sx.lo;
    wy = sy.hi + sy.lo;
  } else {                         // FP32 rows: FP64 dots (the rows carry FP32 rounding anyway)
    wx = 0.0;
    wy = 0.0;
    if (live) {
#pragma unroll
      for (int e = 0; e < CPL; ++e) {
        if (c0 + e < q) {
          wx += (double)cur.xv[e] * pc.cx[e];
          wy += (double)cur.yv[e] * pc.cy[e];
        }
      }
    }
    for (int o = L >> 1; o > 0; o >>= 1) {
      wx += __shfl_xor(wx, o, 64);
      wy += __shfl_xor(wy, o, 64);
    }
  }
  if (!live) return;
  T* xr = X + (int64_t)j * sj + i * si;
  T* yr = Y + (int64_t)j * sj + i * si;
  T xv[CPL], yv[CPL];
  if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      xv[e] = (T)((double)cur.xv[e] - ((ui * pc.cx[e]) / pc.D));
      yv[e] = (T)((double)cur.yv[e] + ((ui * pc.cy[e]) / pc.D));
    }
  } else {
    const double f = ui / pc.D;
#pragma unroll
    for (int e = 0; e < CPL; ++e) {
      xv[e] = (T)((double)cur.xv[e] - f * pc.cx[e]);
      yv[e] = (T)((double)cur.yv[e] + f * pc.cy[e]);
    }
  }
  if (c0 + CPL <= q) {
    vstore_nt<T, CPL>(xr + c0, xv);
    vstore_nt<T, CPL>(yr + c0, yv);
  } else {
#pragma unroll
    for (int e = 0; e < CPL; ++e)
      if (c0 + e < q) {
        xr[c0 + e] = xv[e];
        yr[c0 + e] = yv[e];
      }
  }
}

// Row tiles: a workgroup owns TR = S * (rows per pass) consecutive active rows and walks them pole
// by pole (each pole's TR rows of X and Y are one contiguous run).  w and u of the tile are parked
// in LDS as [row][pole] and written out as one contiguous vertex-major run per array, instead of
// one 8-byte store per (pole, row) p doubles apart.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_maint_rank1_tiles(T* __restrict__ X, T* __restrict__ Y, int64_t sj,
                                                              int64_t si, int p, int64_t nA, int64_t q,
                                                              const T* __restrict__ u, int64_t u_sj, int64_t u_si,
                                                              const double* __restrict__ coef, double* __restrict__ wX,
                                                              double* __restrict__ wY, double* __restrict__ ut, int L,
                                                              int S) {
  extern __shared__ double s_w[];   // [3][TR * p]: w . xi, w . eta, u of (row, pole) at row * p + pole
  Geo g(L);
  const int c0 = g.li * CPL;
  const bool ok = c0 < q;
  const int gpb = kWaves * g.G;
  const int TR = S * gpb;
  const int slot = g.wave * g.G + g.gi;
  const int nk = p * S;             // items of a tile: (pole, pass), pole-major
  const int64_t ntiles = (nA + TR - 1) / TR;
  double* s_x = s_w;
  double* s_y = s_w + (int64_t)TR * p;
  double* s_u = s_w + 2 * (int64_t)TR * p;
  PoleCoef<CPL> pc;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i0 = tile * TR;
    RankRow<T, CPL> cur;
    if (ok && i0 + slot < nA) rank_load<T, CPL>(X, Y, sj, si, q, u, u_sj, u_si, 0, i0 + slot, c0, cur);
    for (int k = 0; k < nk; ++k) {
      const int j = k / S;
      const int r = (k - j * S) * gpb + slot;
      const int64_t i = i0 + r;
      const bool valid = i < nA;
      RankRow<T, CPL> nxt;
      if (k + 1 < nk) {
        const int jn = (k + 1) / S;
        const int64_t in = i0 + ((k + 1) - jn * S) * gpb + slot;
        if (ok && in < nA) rank_load<T, CPL>(X, Y, sj, si, q, u, u_sj, u_si, jn, in, c0, nxt);
      }
      if (j != pc.j) pc.load(coef, q, j, c0);
      const double ui = (ok && valid) ? cur.ui : 0.0;
      double wx, wy;
      rank_item<T, CPL>(X, Y, sj, si, q, L, c0, ok && valid, j, i, cur, pc, ui, wx, wy);
      if (g.li == 0 && valid) {
        s_x[r * p + j] = wx;
        s_y[r * p + j] = wy;
        s_u[r * p + j] = ui;
      }
      cur = nxt;
    }
    __syncthreads();
    const int64_t cnt = min((int64_t)TR, nA - i0) * p;
    for (int64_t t = threadIdx.x; t < cnt; t += kBlock) {
      wX[i0 * p + t] = s_x[t];
      wY[i0 * p + t] = s_y[t];
      ut[i0 * p + t] = s_u[t];
    }
    __syncthreads();
  }
}

// Pole-major walk over all (pole, row) items, for pole counts whose row tile does not fit in LDS:
// w and u go straight to their vertex-major slots.
template <typename T, int CPL>
__global__ __launch_bounds__(kBlock) void k_maint_rank1(T* __restrict__ X, T* __restrict__ Y, int64_t sj, int64_t si,
                                                        int p, int64_t nA, int64_t q, const T* __restrict__ u,
                                                        int64_t u_sj, int64_t u_si, const double* __restrict__ coef,
                                                        double* __restrict__ wX, double* __restrict__ wY,
                                                        double* __restrict__ ut, int L) {
  Geo g(L);
  const int c0 = g.li * CPL;
  const int64_t gpb = (int64_t)kWaves * g.G;
  const int64_t total = (int64_t)p * nA;
  const int64_t step = (int64_t)gridDim.x * gpb;
  const bool ok = c0 < q;
  int64_t item = (int64_t)blockIdx.x * gpb + g.wave * g.G + g.gi;
  RankRow<T, CPL> cur;
  if (item < total && ok) {
    const int j = (int)(item / nA);
    rank_load<T, CPL>(X, Y, sj, si, q, u, u_sj, u_si, j, item - (int64_t)j * nA, c0, cur);
  }
  PoleCoef<CPL> pc;
  for (; item < total; item += step) {
    RankRow<T, CPL> nxt;
    const int64_t nitem = item + step;
    if (nitem < total && ok) {
      const int jn = (int)(nitem / nA);
      rank_load<T, CPL>(X, Y, sj, si, q, u, u_sj, u_si, jn, nitem - (int64_t)jn * nA, c0, nxt);
    }
    const int j = (int)(item / nA);
    const int64_t i = item - (int64_t)j * nA;
    if (j != pc.j) pc.load(coef, q, j, c0);
    const double ui = ok ? cur.ui : 0.0;
    double wx, wy;
    rank_item<T, CPL>(X, Y, sj, si, q, L, c0, ok, j, i, cur, pc, ui, wx, wy);
    if (g.li == 0) {                 // vertex-major: all poles of active row i adjacent
      wX[i * p + j] = wx;
      wY[i * p + j] = wy;
      ut[i * p + j] = ui;
    }
    cur = nxt;
  }
}

constexpr int kMaintLds = 48 * 1024;   // LDS bytes of one k_maint_cache tile

// h/t cache update (polar_gpu.py:246-247).  A workgroup owns a tile of P consecutive pairs and
// every pole.  Phase 1 walks (pair, pole) pole-fastest, so the pair's endpoint runs of ut, wX, wY
// (vertex-major, p doubles per vertex) are read contiguously, and parks cc, dwx, dwy in LDS.
// Phase 2 walks (pole, pair) pair-fastest, so h[j*N + e] and t[j*N + e] are read and written in
// contiguous runs of P (non-temporally: each entry is touched once per call).  Same arithmetic per (pole, pair) as the unfused formula.
__global__ __launch_bounds__(kBlock) void k_maint_cache(int p, int64_t N, int P, const int32_t* __restrict__ ea,
                                                        const int32_t* __restrict__ eb, const double* __restrict__ coef,
                                                        int64_t q, const double* __restrict__ wX,
                                                        const double* __restrict__ wY, const double* __restrict__ ut,
                                                        double* __restrict__ h, double* __restrict__ t) {
  extern __shared__ double lds[];
  const int ld = P + 1;                       // padded pole stride: phase-1 stores spread over banks
  double* s_cc = lds;
  double* s_dx = lds + (int64_t)p * ld;
  double* s_dy = lds + 2 * (int64_t)p * ld;
  const int64_t stride = 2 * q + 3;
  const int items = P * p;
  for (int64_t e0 = (int64_t)blockIdx.x * P; e0 < N; e0 += (int64_t)gridDim.x * P) {
    const int np = (int)((N - e0) < P ? (N - e0) : P);
    for (int it = threadIdx.x; it < np * p; it += blockDim.x) {
      const int k = it / p;
      const int j = it - k * p;
      const int64_t ia = (int64_t)ea[e0 + k] * p + j, ib = (int64_t)eb[e0 + k] * p + j;
      const double D = coef[(int64_t)j * stride + 2 * q];
      s_cc[j * ld + k] = (ut[ia] - ut[ib]) / D;
      s_dx[j * ld + k] = wX[ia] - wX[ib];
      s_dy[j * ld + k] = wY[ia] - wY[ib];
    }
    __syncthreads();
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
      const int j = it / P;
      const int k = it - j * P;
      if (k < np) {
        const double* cj = coef + (int64_t)j * stride;
        const double nxi = cj[2 * q + 1], neta = cj[2 * q + 2];
        const double cc = s_cc[j * ld + k], dwx = s_dx[j * ld + k], dwy = s_dy[j * ld + k];
        const int64_t item = (int64_t)j * N + e0 + k;
        const double hn = (__builtin_nontemporal_load(h + item) - ((2.0 * cc) * dwx)) + ((cc * cc) * nxi);
        const double tn = (__builtin_nontemporal_load(t + item) + ((2.0 * cc) * dwy)) + ((cc * cc) * neta);
        __builtin_nontemporal_store(hn > 0.0 ? hn : 0.0, h + item);
        __builtin_nontemporal_store(tn > 0.0 ? tn : 0.0, t + item);
      }
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void k_rows_axpy(T* __restrict__ F, int64_t ldf, int64_t q, int a, int b, const double* __restrict__ g) {
  for (int64_t c = blockIdx.x * blockDim.x + threadIdx.x; c < q; c += (int64_t)gridDim.x * blockDim.x) {
    F[(int64_t)a * ldf + c] = (T)((double)F[(int64_t)a * ldf + c] + g[c]);
    F[(int64_t)b * ldf + c] = (T)((double)F[(int64_t)b * ldf + c] - g[c]);
  }
}

// ------------------------------------------------------------------------------------------------
// Counter-hash Rademacher generator (splitmix64 finaliser).
__device__ __host__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_rademacher(uint64_t base, const int64_t* __restrict__ keys, int64_t rows,
                                                       int64_t q, double scale, T* __restrict__ out, int64_t ld) {
  const int64_t total = rows * q;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / q, c = idx - i * q;
    const uint64_t key = keys ? (uint64_t)keys[i] : (uint64_t)i;
    const uint64_t hsh = splitmix64(base + key * (uint64_t)q + (uint64_t)c);
    out[i * ld + c] = (T)((hsh >> 63) ? -scale : scale);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_edge_sketch(polar_op op, uint64_t base, int64_t q, double scale,
                                                        T* __restrict__ F, int64_t ldf) {
  // one thread per (row, column); loops over the row's neighbours in CSR order
  const int64_t total = op.n * q;
  const T* wt = (const T*)op.weights;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / q, c = idx - i * q;
    double acc = 0.0;
    for (int k = op.indptr[i]; k < op.indptr[i + 1]; ++k) {
      const int64_t j = op.indices[k];
      const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
      const uint64_t hsh = splitmix64(base + (uint64_t)(lo * op.n + hi) * (uint64_t)q + (uint64_t)c);
      const double w = op.weighted ? sqrt((double)wt[k]) : 1.0;
      const double sgn = (i < j) ? 1.0 : -1.0;
      acc += sgn * w * ((hsh >> 63) ? -scale : scale);
    }
    F[i * ldf + c] = (T)acc;
  }
}

static int grid_for(int64_t items, int per_block, int cap = 65536) {
  return (int)std::min<int64_t>(cap, std::max<int64_t>(1, (items + per_block - 1) / per_block));
}

static int pick_cpl_rows(int dtype, int64_t q, const void* base, int64_t sj, int64_t si) {
  const int esz = elem_size(dtype);
  for (int cpl = 16 / esz; cpl > 1; cpl /= 2) {
    if (q % cpl == 0 && ((uintptr_t)base % (cpl * esz)) == 0 && (sj * esz) % (cpl * esz) == 0 &&
        (si * esz) % (cpl * esz) == 0 && (q + cpl - 1) / cpl <= 64)
      return cpl;
  }
  return 1;
}

constexpr int kRankLds = 48 * 1024;   // LDS bytes of one k_maint_rank1_tiles row tile

// Row tiles of S passes (S <= 4) when one pass of rows fits the LDS budget, else the pole-major walk.
template <typename T, int CPL>
static void launch_rank1(T* X, T* Y, int64_t sj, int64_t si, int p, int64_t nA, int64_t q, const T* u,
                         int64_t u_sj, int64_t u_si, const double* coef, double* wX, double* wY, double* ut, int L,
                         hipStream_t st) {
  const int gpb = kWaves * (64 / L);
  const int64_t S = std::min<int64_t>(4, kRankLds / (24 * (int64_t)gpb * p));
  if (S >= 1) {
    const int64_t TR = S * gpb;
    const int grid = grid_for((nA + TR - 1) / TR, 1);
    hipLaunchKernelGGL((k_maint_rank1_tiles<T, CPL>), dim3(grid), dim3(kBlock), (size_t)(24 * TR * p), st, X, Y, sj,
                       si, p, nA, q, u, u_sj, u_si, coef, wX, wY, ut, L, (int)S);
  } else {
    hipLaunchKernelGGL((k_maint_rank1<T, CPL>), dim3(grid_for((int64_t)p * nA, gpb)), dim3(kBlock), 0, st, X, Y, sj,
                       si, p, nA, q, u, u_sj, u_si, coef, wX, wY, ut, L);
  }
}

}  // namespace polar

using namespace polar;

extern "C" {

int polar_pair_sqnorm(int32_t dtype, const void* X, int64_t x_sj, int64_t x_si, int32_t p, int64_t q,
                      const int32_t* ea, const int32_t* eb, int64_t N, double* out, void* stream) {
  if (!dtype_ok(dtype) || !X || !ea || !eb || !out || p <= 0 || q <= 0 || N < 0)
    return set_error(POLAR_E_ARG, "bad pair_sqnorm args");
  if (N == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int cpl = pick_cpl_rows(dtype, q, X, x_sj, x_si);
  int L = lanes_for(q, cpl);
  if (L > 64) {
    cpl = (dtype == POLAR_F64) ? 2 : 4;
    L = 64;
    if ((q + cpl - 1) / cpl > 64) return set_error(POLAR_E_ARG, "q=%lld too wide", (long long)q);
    return set_error(POLAR_E_ARG, "q=%lld needs aligned 16-byte rows", (long long)q);
  }
  const int gpb = kWaves * (64 / L);
  const int grid = grid_for((int64_t)p * N, gpb);
  if (dtype == POLAR_F64) {
    if (cpl == 2) hipLaunchKernelGGL((k_pair_sqnorm<double, 2>), dim3(grid), dim3(kBlock), 0, st, (const double*)X, x_sj, x_si, p, q, ea, eb, N, out, L);
    else hipLaunchKernelGGL((k_pair_sqnorm<double, 1>), dim3(grid), dim3(kBlock), 0, st, (const double*)X, x_sj, x_si, p, q, ea, eb, N, out, L);
  } else {
    if (cpl == 4) hipLaunchKernelGGL((k_pair_sqnorm<float, 4>), dim3(grid), dim3(kBlock), 0, st, (const float*)X, x_sj, x_si, p, q, ea, eb, N, out, L);
    else if (cpl == 2) hipLaunchKernelGGL((k_pair_sqnorm<float, 2>), dim3(grid), dim3(kBlock), 0, st, (const float*)X, x_sj, x_si, p, q, ea, eb, N, out, L);
    else hipLaunchKernelGGL((k_pair_sqnorm<float, 1>), dim3(grid), dim3(kBlock), 0, st, (const float*)X, x_sj, x_si, p, q, ea, eb, N, out, L);
  }
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_pair_wsum(int32_t dtype, const void* X, int64_t x_sj, int64_t x_si, int32_t p, int64_t q,
                    const int32_t* ea, const int32_t* eb, int64_t E, const double* om, double* out,
                    int32_t accumulate, void* stream) {
  if (!dtype_ok(dtype) || !X || !ea || !eb || !om || !out || p <= 0 || q <= 0 || E < 0)
    return set_error(POLAR_E_ARG, "bad pair_wsum args");
  if (E == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int cpl = pick_cpl_rows(dtype, q, X, x_sj, x_si);
  const int L = lanes_for(q, cpl);
  if (L > 64) return set_error(POLAR_E_ARG, "q=%lld too wide for pair_wsum at this alignment", (long long)q);
  const int grid = grid_for(E, kWaves * (64 / L));
  if (dtype == POLAR_F64) {
    if (cpl == 2) hipLaunchKernelGGL((k_pair_wsum<double, 2>), dim3(grid), dim3(kBlock), 0, st, (const double*)X, x_sj, x_si, p, q, ea, eb, E, om, out, accumulate, L);
    else hipLaunchKernelGGL((k_pair_wsum<double, 1>), dim3(grid), dim3(kBlock), 0, st, (const double*)X, x_sj, x_si, p, q, ea, eb, E, om, out, accumulate, L);
  } else {
    if (cpl == 4) hipLaunchKernelGGL((k_pair_wsum<float, 4>), dim3(grid), dim3(kBlock), 0, st, (const float*)X, x_sj, x_si, p, q, ea, eb, E, om, out, accumulate, L);
    else if (cpl == 2) hipLaunchKernelGGL((k_pair_wsum<float, 2>), dim3(grid), dim3(kBlock), 0, st, (const float*)X, x_sj, x_si, p, q, ea, eb, E, om, out, accumulate, L);
    else hipLaunchKernelGGL((k_pair_wsum<float, 1>), dim3(grid), dim3(kBlock), 0, st, (const float*)X, x_sj, x_si, p, q, ea, eb, E, om, out, accumulate, L);
  }
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_proxy(int32_t obj, const double* h, const double* t, const double* tau, const double* om, int32_t p,
                int64_t N, double sigma, const uint8_t* dead, double* score, void* stream) {
  if (!h || !t || !tau || !om || !score || p <= 0 || N < 0) return set_error(POLAR_E_ARG, "bad proxy args");
  if (N == 0) return 0;
  hipLaunchKernelGGL(k_proxy, dim3(grid_for(N, kBlock, 8192)), dim3(kBlock), 0, (hipStream_t)stream, obj, h, t, tau,
                     om, p, N, sigma, dead, score);
  POLAR_LAUNCH_CHECK();
  return 0;
}

int64_t polar_select_ws_bytes(int64_t N) {
  (void)N;
  return (int64_t)kMaxGrid * 16 + 64 * 8 + 256 + 512;
}

int polar_select_topb(const double* score, int64_t N, int32_t Br, int32_t* out, int32_t* host_out, void* ws,
                      int64_t ws_bytes, void* stream) {
  if (!score || !out || Br <= 0 || Br > 64 || Br > N || !ws) return set_error(POLAR_E_ARG, "bad select args (Br=%d N=%lld)", Br, (long long)N);
  if (ws_bytes < polar_select_ws_bytes(N)) return set_error(POLAR_E_ARG, "select workspace too small");
  hipStream_t st = (hipStream_t)stream;
  char* w = (char*)ws;
  double* bs = (double*)w;
  int64_t* bi = (int64_t*)(w + kMaxGrid * 8);
  double* sel_s = (double*)(w + kMaxGrid * 16);
  const int grid = grid_for(N, kBlock, kBlock);   // <= 256 workgroup results per pass
  for (int b = 0; b < Br; ++b) {
    hipLaunchKernelGGL(k_select_blocks, dim3(grid), dim3(kBlock), 0, st, score, N, b, (const double*)sel_s,
                       (const int32_t*)out, bs, bi);
    POLAR_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_select_final, dim3(1), dim3(kBlock), 0, st, b, grid, (const double*)bs, (const int64_t*)bi,
                       sel_s, out);
    POLAR_LAUNCH_CHECK();
  }
  if (host_out) {
    POLAR_HIP(hipMemcpyAsync(host_out, out, sizeof(int32_t) * Br, hipMemcpyDeviceToHost, st));
    POLAR_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

int polar_pair_rhs(int32_t dtype, int64_t n, int32_t Br, const int32_t* a, const int32_t* b, void* W, void* stream) {
  if (!dtype_ok(dtype) || !a || !b || !W || Br <= 0 || n <= 0) return set_error(POLAR_E_ARG, "bad pair_rhs args");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = zero_async(W, (size_t)n * Br * elem_size(dtype), st)) return rc;
  if (dtype == POLAR_F64) hipLaunchKernelGGL(k_pair_rhs<double>, dim3((Br + 63) / 64), dim3(64), 0, st, Br, a, b, (double*)W);
  else hipLaunchKernelGGL(k_pair_rhs<float>, dim3((Br + 63) / 64), dim3(64), 0, st, Br, a, b, (float*)W);
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_rerank_pick(int32_t obj, int32_t dtype, const void* U, int64_t u_sj, int64_t u_si, int32_t p, int32_t Br,
                      const int32_t* sa, const int32_t* sb, const double* hh, const double* om, double* g,
                      int32_t* host_winner, double* host_value, void* stream) {
  if (!dtype_ok(dtype) || !U || !sa || !sb || !om || !g || Br <= 0 || p <= 0)
    return set_error(POLAR_E_ARG, "bad rerank args");
  if (obj != 2 && !hh) return set_error(POLAR_E_ARG, "rerank needs tracked norms");
  hipStream_t st = (hipStream_t)stream;
  double* res = g + Br;  // caller provides Br + 2 doubles
  if (dtype == POLAR_F64)
    hipLaunchKernelGGL(k_rerank<double>, dim3(1), dim3(kBlock), 0, st, obj, (const double*)U, u_sj, u_si, p, Br, sa, sb, hh, om, g, res);
  else
    hipLaunchKernelGGL(k_rerank<float>, dim3(1), dim3(kBlock), 0, st, obj, (const float*)U, u_sj, u_si, p, Br, sa, sb, hh, om, g, res);
  POLAR_LAUNCH_CHECK();
  if (!host_winner && !host_value) return 0;     // the caller reads g[Br], g[Br + 1] itself
  double hr[2];
  POLAR_HIP(hipMemcpyAsync(hr, res, sizeof(hr), hipMemcpyDeviceToHost, st));
  POLAR_HIP(hipStreamSynchronize(st));
  if (host_winner) *host_winner = (int32_t)hr[0];
  if (host_value) *host_value = hr[1];
  return 0;
}

int polar_maint_prep(int32_t dtype, const void* X, const void* Y, int64_t sj, int64_t si, int32_t p, int64_t q,
                     const void* u, int64_t u_sj, int64_t u_si, int32_t aa, int32_t bb, const double* gc,
                     double* coef, void* stream) {
  if (!dtype_ok(dtype) || !X || !Y || !u || !gc || !coef || p <= 0 || q <= 0)
    return set_error(POLAR_E_ARG, "bad maint_prep args");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == POLAR_F64)
    hipLaunchKernelGGL(k_maint_prep<double>, dim3(p), dim3(kBlock), 0, st, (const double*)X, (const double*)Y, sj, si, q, (const double*)u, u_sj, u_si, aa, bb, gc, coef);
  else
    hipLaunchKernelGGL(k_maint_prep<float>, dim3(p), dim3(kBlock), 0, st, (const float*)X, (const float*)Y, sj, si, q, (const float*)u, u_sj, u_si, aa, bb, gc, coef);
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_maint_rank1(int32_t dtype, void* X, void* Y, int64_t sj, int64_t si, int32_t p, int64_t nA, int64_t q,
                      const void* u, int64_t u_sj, int64_t u_si, const double* coef, double* wX, double* wY,
                      double* ut, void* stream) {
  if (!dtype_ok(dtype) || !X || !Y || !u || !coef || !wX || !wY || !ut || p <= 0 || q <= 0 || nA < 0)
    return set_error(POLAR_E_ARG, "bad maint_rank1 args");
  if (nA == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int cpl = std::min(pick_cpl_rows(dtype, q, X, sj, si), pick_cpl_rows(dtype, q, Y, sj, si));
  int L = lanes_for(q, cpl);
  if (L > 64) return set_error(POLAR_E_ARG, "q=%lld too wide for maint_rank1 at this alignment", (long long)q);
  if (dtype == POLAR_F64) {
    if (cpl == 2) launch_rank1<double, 2>((double*)X, (double*)Y, sj, si, p, nA, q, (const double*)u, u_sj, u_si, coef, wX, wY, ut, L, st);
    else launch_rank1<double, 1>((double*)X, (double*)Y, sj, si, p, nA, q, (const double*)u, u_sj, u_si, coef, wX, wY, ut, L, st);
  } else {
    if (cpl == 4) launch_rank1<float, 4>((float*)X, (float*)Y, sj, si, p, nA, q, (const float*)u, u_sj, u_si, coef, wX, wY, ut, L, st);
    else if (cpl == 2) launch_rank1<float, 2>((float*)X, (float*)Y, sj, si, p, nA, q, (const float*)u, u_sj, u_si, coef, wX, wY, ut, L, st);
    else launch_rank1<float, 1>((float*)X, (float*)Y, sj, si, p, nA, q, (const float*)u, u_sj, u_si, coef, wX, wY, ut, L, st);
  }
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_maint_cache(int32_t p, int64_t N, const int32_t* ea, const int32_t* eb, const double* coef, int64_t q,
                      const double* wX, const double* wY, const double* ut, double* h, double* t, void* stream) {
  if (!ea || !eb || !coef || !wX || !wY || !ut || !h || !t || p <= 0 || q <= 0 || N < 0)
    return set_error(POLAR_E_ARG, "bad maint_cache args");
  if (N == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  // Pairs per tile: 3 LDS doubles per (pole, pair) within 48 KiB, padded stride P + 1.
  const int P = std::min<int64_t>(64, (int64_t)kMaintLds / (24 * (int64_t)p) - 1);
  if (P < 1) return set_error(POLAR_E_ARG, "maint_cache: p=%d poles exceed the LDS tile", p);
  const size_t lds = (size_t)24 * p * (P + 1);
  const int grid = grid_for((N + P - 1) / P, 1);
  hipLaunchKernelGGL(k_maint_cache, dim3(grid), dim3(kBlock), lds, st, p, N, P, ea, eb, coef, q, wX, wY, ut, h, t);
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_rows_axpy(int32_t dtype, void* F, int64_t ldf, int64_t q, int32_t a, int32_t b, const double* g, void* stream) {
  if (!dtype_ok(dtype) || !F || !g || q <= 0 || ldf < q || a < 0 || b < 0) return set_error(POLAR_E_ARG, "bad rows_axpy args");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == POLAR_F64) hipLaunchKernelGGL(k_rows_axpy<double>, dim3(1), dim3(256), 0, st, (double*)F, ldf, q, a, b, g);
  else hipLaunchKernelGGL(k_rows_axpy<float>, dim3(1), dim3(256), 0, st, (float*)F, ldf, q, a, b, g);
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_rademacher(int32_t dtype, uint64_t seed, int32_t stream_id, const int64_t* keys, int64_t rows, int64_t q,
                     double scale, void* out, int64_t ld, void* stream) {
  if (!dtype_ok(dtype) || !out || rows < 0 || q <= 0 || ld < q) return set_error(POLAR_E_ARG, "bad rademacher args");
  if (rows == 0) return 0;
  const uint64_t base = splitmix64(seed * 4 + (uint64_t)stream_id);
  hipStream_t st = (hipStream_t)stream;
  const int grid = grid_for(rows * q, kBlock);
  if (dtype == POLAR_F64) hipLaunchKernelGGL(k_rademacher<double>, dim3(grid), dim3(kBlock), 0, st, base, keys, rows, q, scale, (double*)out, ld);
  else hipLaunchKernelGGL(k_rademacher<float>, dim3(grid), dim3(kBlock), 0, st, base, keys, rows, q, scale, (float*)out, ld);
  POLAR_LAUNCH_CHECK();
  return 0;
}

int polar_edge_sketch(const polar_op* op, uint64_t seed, int64_t q, double scale, void* F, int64_t ldf, void* stream) {
  if (!op || !dtype_ok(op->dtype) || op->n <= 0 || !op->indptr || (op->weighted && !op->weights) || !F || q <= 0 ||
      ldf < q)
    return set_error(POLAR_E_ARG, "bad edge_sketch args");
  const uint64_t base = splitmix64(seed * 4 + 1);
  hipStream_t st = (hipStream_t)stream;
  const int grid = grid_for(op->n * q, kBlock);
  if (op->dtype == POLAR_F64) hipLaunchKernelGGL(k_edge_sketch<double>, dim3(grid), dim3(kBlock), 0, st, *op, base, q, scale, (double*)F, ldf);
  else hipLaunchKernelGGL(k_edge_sketch<float>, dim3(grid), dim3(kBlock), 0, st, *op, base, q, scale, (float*)F, ldf);
  POLAR_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
