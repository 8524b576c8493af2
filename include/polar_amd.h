/* polar_amd.h -- C-ABI of the MI355X (gfx950) pole-shared Krylov core.
 *
 * Every entry point takes raw device pointers owned by the caller, sizes as int64_t, and an
 * explicit hipStream_t (passed as void*).  The library allocates no device memory: temporaries
 * live in a caller-provided workspace whose size is returned by the matching *_ws_bytes query.
 * Host-side state: polar_msolve keeps, per calling thread, a small pinned buffer for its
 * convergence checks and (with POLAR_GRAPH) a capture stream; polar_release() frees them.
 * Return value: 0 on success, a positive hipError_t, or a negative POLAR_E_* code; the text of
 * the last error is returned by polar_last_error().  No entry point synchronises the stream
 * except polar_msolve, which reads the convergence flags of each batch of 8 iterations (as the
 * reference's host check at realworld/polar_gpu.py:126) while the next batch is already queued,
 * and polar_select_topb / polar_rerank_pick when asked for host copies of their results (the
 * shortlist and winner the reference copies back, polar_gpu.py:221, :236).
 * Memory that must start at zero is cleared by a fill kernel, never by hipMemsetAsync, so a
 * caller that captures an entry point without a host sync into a HIP graph records no memset
 * node.  Sizes, pointers and dtype codes are checked on the host before anything is launched
 * (POLAR_E_ARG).
 *
 * Reference interfaces replaced (chenty2333/POLAR):
 *   realworld/spmm.py:3-14          csr_spmm(n, R, ptr, idx, val, X, Y)   -> polar_csr_spmm
 *   realworld/spmm.py:18-25         spmm(G.M0, X) on the graph Laplacian  -> polar_spmm
 *   realworld/polar_gpu.py:63-80    Op.__call__(x, shift)                  -> polar_spmm (patch)
 *   realworld/polar_gpu.py:88-133   msolve(op, W, tau, active, tol, ...)   -> polar_msolve
 *   realworld/polar_gpu.py:205      bank gather-reduce h, t                -> polar_pair_sqnorm
 *   realworld/polar_gpu.py:208-211  proxy(h, t) + dead mask (:219)         -> polar_proxy
 *   realworld/polar_gpu.py:221      argsort(-sc)[:Br]                      -> polar_select_topb
 *   realworld/polar_gpu.py:225-236  Wb build, bu gather, re-rank argmax    -> polar_pair_rhs,
 *                                                                             polar_rerank_pick
 *   realworld/polar_gpu.py:243-249  Sherman-Morrison maintenance          -> polar_maint_prep,
 *                                                                             polar_maint_rank1,
 *                                                                             polar_maint_cache
 *   realworld/polar_gpu.py:250      F[a] += Gc[e]; F[b] -= Gc[e]           -> polar_rows_axpy
 *   realworld/polar_gpu.py:148-165  sketches (counter-hash Rademacher kind) -> polar_rademacher
 */
#ifndef POLAR_AMD_H
#define POLAR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POLAR_ABI_VERSION 5

enum polar_dtype { POLAR_F64 = 0, POLAR_F32 = 1 };

enum polar_status {
  POLAR_OK = 0,
  POLAR_E_ITERCAP = -2,   /* multi-shift CG iteration cap (polar_gpu.py:128-129) */
  POLAR_E_ARG = -3,       /* invalid argument (shape, dtype, alignment, workspace size) */
};

/* msolve flags */
#define POLAR_TRACK 1     /* track ||u_j||^2 with three scalars per pole (polar_gpu.py:114-117) */
#define POLAR_TRACE 2     /* track w^T (M + tau_j)^{-1} w with two scalars per pole */
#define POLAR_NOFREEZE 4  /* never freeze converged poles (synthetic/scalability/benchmark.py:72) */
#define POLAR_GRAPH 8     /* replay the CG iterations as hipGraphs of 8 iterations (same results) */
#define POLAR_BLOCK8 16   /* shifted x/p updates once per 8 iterations from the stored residuals:
                             same CG scalars and iteration counts, X within rounding of the
                             per-iteration updates, about 8x fewer passes over the shifted blocks */
#define POLAR_REPLACE 32  /* residual replacement every 24 iterations: r = w - (M + tau0) x0 with the
                             seed iterate x0 kept in FP64 and the product in FP64 (for FP32 vectors) */
#define POLAR_HISTORY 64  /* the shifted blocks are not formed: the residual of every iteration and
                             every pole's per-iteration coefficients stay in the workspace (maxit
                             iterations at most, see polar_msolve_history_ws_bytes), and
                             polar_history_rows / polar_history_cols form X_j afterwards on chosen
                             rows of every column, or on every active row of chosen columns.  One
                             segment (column groups allowed); no column compaction; implies
                             POLAR_BLOCK8's window coefficients; Xout may be NULL; POLAR_GRAPH is
                             ignored.  No counterpart in the reference: its re-rank solve forms
                             every shortlisted column's X (polar_gpu.py:228) and then uses the
                             winner's (:237) and two rows of the others (:230). */

/* Operator y = (diag + shift) x - A_w x + patch(x): the graph Laplacian M0 = diag(deg+sigma) - A
 * of realworld/polar_gpu.py:47-55 stored as diagonal + off-diagonal adjacency (value-free when
 * every weight is 1), plus the unit-weight added edges of polar_gpu.py:68-79 as a row patch.
 * Row sums run in adjacency order; rows longer than `chunk` entries are split into chunks of
 * `chunk` entries whose partial sums are added in chunk order (the schedule is built by the host).
 * POLAR_CHUNK is the default chunk length. */
#define POLAR_CHUNK 64
typedef struct polar_op {
  int64_t n;
  int32_t dtype;              /* polar_dtype of weights, diag and all vectors */
  int32_t weighted;           /* 0: all off-diagonal weights are 1 and `weights` is NULL */
  const int32_t* indptr;      /* [n+1] */
  const int32_t* indices;     /* [nnz]  off-diagonal columns */
  const void* weights;        /* [nnz]  off-diagonal weights (positive), or NULL */
  const void* diag;           /* [n]    weighted degree + sigma */
  int32_t n_long;             /* rows with degree > chunk */
  int32_t n_extra;            /* entries of `extra` (= n_partial: every chunk of every long row) */
  int32_t n_partial;          /* total chunks of long rows */
  int32_t chunk;              /* long-row chunk length (>= 8) */
  const int32_t* long_rows;   /* [n_long] ascending row ids */
  const int32_t* long_base;   /* [n_long] first partial slot of each long row */
  const int32_t* extra;       /* [n_extra][2]: (long index, chunk index) */
  int32_t n_patch;            /* rows touched by added edges */
  int32_t cpl_max;            /* tuning: max columns per lane (1, 2, 4; 0 = widest 16-byte load) */
  const int32_t* patch_rows;  /* [n_patch] ascending */
  const int32_t* patch_ptr;   /* [n_patch+1] */
  const int32_t* patch_cols;  /* [patch_ptr[n_patch]] */
  int32_t grid_max;           /* tuning: block budget of the kernels with partial sums (0 = default:
                                 1024 blocks, 16384 for the operator kernel on >= 2^20 rows) */
  int32_t run_rows;           /* tuning: consecutive rows per work unit of the operator kernel (their
                                 gathers share batches; capped below the lanes per row); 0 or 1 =
                                 one row per unit */
  const int32_t* patch_grp;   /* [patch_ptr[n_patch]] column group of each patch entry, or NULL
                                 (every entry applies to every column) */
} polar_op;

/* Segments of a block-diagonal operator (the batched form: many graphs in one operator, each on
 * its own row range, with its own poles and CG scalars).  Host arrays of S+1 ascending offsets:
 * rows, entries of op.extra (long-row chunks), long rows (op.long_rows), active rows.  Segment g
 * owns rows [rows[g], rows[g+1]) and its long rows, chunks and active rows lie in its ranges. */
typedef struct polar_segs {
  int32_t S;
  int32_t G;                  /* column groups (S == 1): each group of columns is its own system on
                                 the one graph, with poles tau[g*p..], seed shift tau[g*p] and the
                                 patch entries tagged g; 0 or 1 = none */
  const int64_t* rows;
  const int64_t* extra;
  const int64_t* longs;
  const int64_t* active;
  const int32_t* col_group;   /* host int32[R]: group of each column of this call (G > 1) */
} polar_segs;

int polar_abi_version(void);
const char* polar_last_error(void);

/* Frees the calling thread's host-side solver resources (the pinned convergence-check buffer and
 * the hipGraph capture stream); the next polar_msolve call recreates them.  Returns 0 or the
 * hipError_t of a failed free.  No counterpart in the reference (CuPy owns its pools). */
int polar_release(void);

/* Live profiling of the CG iteration kernels with HIP events (no counterpart in the reference;
 * used by the benchmark).  polar_prof_set(1) resets and enables it for every class;
 * polar_prof_set(2) times class 0 only (two events per iteration instead of seven);
 * polar_prof_set(3) times class 0 on one iteration in every 8 (the position rotates from one
 * 8-iteration batch to the next), so the timed work is barely perturbed; 0 disables it.
 * Classes 0..5: k_opmul (operator product + d.v partials), k_longfin (long-row chunk combine),
 * k_alpha, k_resid (residual update + |r|^2 partials), k_scalars (per-pole recurrences),
 * k_update (direction and shifted x/p updates); per class the summed event time (ms), the timed
 * launches that did work, and their algorithmic HBM bytes.  Class 6: solves, and (bytes) the
 * number of live (pole, column) pairs summed over iterations.  Class 7 (launches): every
 * iteration that did work, timed or not.  Any other mode is rejected (POLAR_E_ARG). */
int polar_prof_set(int32_t on);
int polar_prof_get(int32_t cls, double* ms, int64_t* launches, double* bytes);

/* Zero-allocation form of polar_spmm, for callers that cannot allocate device memory or pass a
 * workspace (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same product and kernels as polar_spmm; the long-row chunk
 * partials live in a file-scope device array of polar_spmm_static_max_bytes() bytes (32 MiB) that
 * each GPU holds until the process ends.  Rejects (POLAR_E_ARG) an operator and width whose
 * partials need more (polar_spmm_ws_bytes(op, R) > the maximum).  Calls no allocator.  Calls that
 * overlap on one GPU (other streams or threads) share the array and race: serialise them. */
int64_t polar_spmm_static_max_bytes(void);
int polar_spmm_static(const polar_op* op, double shift, const void* X, int64_t ldx, void* Y, int64_t ldy,
                      int64_t R, void* stream);

/* General CSR product Y[:, :R] = A X[:, :R] (realworld/spmm.py:3-14, csr_spmm(n, R, ptr, idx, val,
 * X, Y)): A is n rows of CSR with int32 indptr[n+1] and indices[nnz] and values[nnz] of `dtype`
 * (any values, diagonal stored like the reference's G.M0), X and Y row-major with leading dims
 * ldx, ldy (the reference's kernel has ldx = ldy = R).  Each row is summed in entry order; rows
 * with no entries give 0.  Any R >= 1 (the reference hands R < 8 to cuSPARSE, the same product).
 * Allocates nothing, needs no workspace; asynchronous on `stream`. */
int polar_csr_spmm(int32_t dtype, int64_t n, int64_t R, const int32_t* indptr, const int32_t* indices,
                   const void* values, const void* X, int64_t ldx, void* Y, int64_t ldy, void* stream);

/* Y[:, :R] = (M0 + shift I + sum_S b b^T) X[:, :R]; X, Y row-major with leading dims ldx, ldy:
 * the Laplacian operator of polar_op (value-free adjacency + diagonal + added-edge patch), the
 * form every solve uses (realworld/polar_gpu.py:71-80 Op.__call__ through spmm.py:18-25). */
int64_t polar_spmm_ws_bytes(const polar_op* op, int64_t R);
int polar_spmm(const polar_op* op, double shift, const void* X, int64_t ldx, void* Y, int64_t ldy,
               int64_t R, void* ws, int64_t ws_bytes, void* stream);

/* Block multi-shift CG for (M + tau_j I) U = W[:, :R] on all poles (realworld/polar_gpu.py:88-133).
 * tau: device double[p], ascending (tau[0] is the seed shift).  active: device int32[nA].
 * Xout: element (j, i, c) at Xout + j*x_sj + i*x_si + c (dtype of op).
 * h_out (TRACK), te_out (TRACE): device double[p*R] or NULL.  iters_out: device int64[R].
 * Returns POLAR_E_ITERCAP when a column is still active after the last 8-iteration check within
 * maxit iterations (the reference's for/else at polar_gpu.py:100-129). */
int64_t polar_msolve_ws_bytes(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags);
int polar_msolve(const polar_op* op, const void* W, int64_t ldw, int64_t R, const double* tau,
                 int32_t p, const int32_t* active, int64_t nA, double tol, int64_t maxit,
                 int32_t flags, void* Xout, int64_t x_sj, int64_t x_si, double* h_out,
                 double* te_out, int64_t* iters_out, void* ws, int64_t ws_bytes, void* stream);

/* Batched form of polar_msolve: S independent systems (M_g + tau[g*p + j] I) U = W[rows of g, :R]
 * on the segments of one block-diagonal operator, advanced together -- one launch per kernel for
 * all segments, each (segment, column) with its own CG scalars, finished segments skipped and
 * their blocks given to the others.  Segment g's results equal those of polar_msolve on graph g
 * alone.  With segs->G > 1 (one segment), the columns are G groups instead: column c solves with
 * the poles of group col_group[c] and the patch entries of that group (several s values, each
 * with its own selected edges, on one graph: one pass over the adjacency for all of them); tau is
 * then double[G][p] and h_out, te_out are [p][R].  The reference runs graphs one at a time (realworld/run_real.py:114-133 loops over
 * graphs calling polar_gpu.msolve, polar_gpu.py:88-133); this entry point has no counterpart
 * there.  tau: device double[S][p].  active: device int32[nA], nA = segs->active[S], union row
 * ids.  Xout (j, i, c) at Xout + j*x_sj + i*x_si + c over the nA active rows.  h_out, te_out:
 * device double[p][S][R] or NULL; iters_out: device int64[S][R]. */
int64_t polar_msolve_batched_ws_bytes(const polar_op* op, const polar_segs* segs, int64_t R, int32_t p,
                                      int32_t flags);
int polar_msolve_batched(const polar_op* op, const polar_segs* segs, const void* W, int64_t ldw, int64_t R,
                         const double* tau, int32_t p, const int32_t* active, double tol, int64_t maxit,
                         int32_t flags, void* Xout, int64_t x_sj, int64_t x_si, double* h_out,
                         double* te_out, int64_t* iters_out, void* ws, int64_t ws_bytes, void* stream);

/* Workspace of a POLAR_HISTORY solve of R columns (polar_msolve, or polar_msolve_batched with one
 * segment) that keeps up to tcap iterations; pass maxit = tcap to the solve (a solve that needs
 * more returns POLAR_E_ITERCAP and can be repeated without POLAR_HISTORY). */
int64_t polar_msolve_history_ws_bytes(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags,
                                      int64_t tcap);

/* After a POLAR_HISTORY solve with this workspace (same op, R, p, nA, flags, tcap): the shifted
 * solutions on K rows of every column, E[j][k][c] = X_j[rows[k*R + c], c] (rows: device int32
 * absolute row ids in [0, n), read on the device; E: device, the op's dtype, p*K*R values).
 * Asynchronous; reads the history, so call it before the workspace is reused.  For the re-rank this is
 * U[j, a_c, c] and U[j, b_c, c] of realworld/polar_gpu.py:230 (K = 2, rows = the endpoints). */
int polar_history_rows(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags, int64_t tcap,
                       void* ws, int64_t ws_bytes, int32_t K, const int32_t* rows, void* E, void* stream);

/* After a POLAR_HISTORY solve: X_j on every active row for the C columns cols[] (device int32),
 * X[(i*p + j)*nA + a] = X_j[active[a], cols[i]] (device, the op's dtype): the winner's u of
 * realworld/polar_gpu.py:237.  At most 4 columns per call and 64 poles; cols in [0, R) and active
 * the solve's active rows, both read on the device.  Asynchronous. */
int polar_history_cols(const polar_op* op, int64_t R, int32_t p, int64_t nA, int32_t flags, int64_t tcap,
                       void* ws, int64_t ws_bytes, int32_t C, const int32_t* cols, const int32_t* active,
                       void* X, void* stream);

/* out[j*N + e] = sum_c (X[j, ea[e], c] - X[j, eb[e], c])^2, double (polar_gpu.py:205). */
int polar_pair_sqnorm(int32_t dtype, const void* X, int64_t x_sj, int64_t x_si, int32_t p,
                      int64_t q, const int32_t* ea, const int32_t* eb, int64_t N, double* out,
                      void* stream);

/* out[e] (+)= sum_j om_j sum_c (X[j, ea[e], c] - X[j, eb[e], c])^2 over E index pairs (rows of X).
 * With X_j = (M + tau_j)^{-1} Z on all rows this is minus the derivative of the sketched trace
 * sum_j om_j tr(Z^T (M + tau_j)^{-1} Z) with respect to the weight of pair e (the adjoint form;
 * no counterpart kernel in the reference, whose h cache at polar_gpu.py:205 holds the per-pole
 * terms for candidate pairs).  accumulate != 0 adds to out (column-chunked streaming). */
int polar_pair_wsum(int32_t dtype, const void* X, int64_t x_sj, int64_t x_si, int32_t p, int64_t q,
                    const int32_t* ea, const int32_t* eb, int64_t E, const double* om, double* out,
                    int32_t accumulate, void* stream);

/* score[e] = sum_j om_j h/(1 + t + (sigma + tau_j) h)  (obj 0 = frac/er), or
 * log1p(t[0,e] + sigma h[0,e]) (obj 2 = logdet); dead[e] != 0 -> -inf (polar_gpu.py:208-219).
 * obj 3: the first-order gain sum_j om_j h[j,e] = -d(sketched trace)/dw_e at w_e = 0 (the
 * adjoint form's gradient shortlist; no counterpart in the reference). */
int polar_proxy(int32_t obj, const double* h, const double* t, const double* tau, const double* om,
                int32_t p, int64_t N, double sigma, const uint8_t* dead, double* score, void* stream);

/* First Br indices of the stable descending order of score (ties: lower index first), i.e.
 * argsort(-score, stable)[:Br] (polar_gpu.py:221).  Writes device int32[Br] to out and, when
 * host_out is not NULL, copies them to host_out (synchronising the stream).
 * ws: polar_select_ws_bytes(N) bytes. */
int64_t polar_select_ws_bytes(int64_t N);
int polar_select_topb(const double* score, int64_t N, int32_t Br, int32_t* out, int32_t* host_out,
                      void* ws, int64_t ws_bytes, void* stream);

/* W[:, c] = e_a - e_b for pairs (a[c], b[c]), c < Br; W is n x Br row-major, zeroed first. */
int polar_pair_rhs(int32_t dtype, int64_t n, int32_t Br, const int32_t* a, const int32_t* b,
                   void* W, void* stream);

/* Re-rank (polar_gpu.py:229-236): bu[j,c] = U[j, sa[c], c] - U[j, sb[c], c];
 * g[c] = sum_j om_j hh[j,c]/(1+bu[j,c]) (obj 0) or log1p(bu[0,c]) (obj 2);
 * winner = first argmax.  Writes g (device double[Br + 2]: the scores, then the winner as a double
 * and its score) and copies (winner, g[winner]) to host_winner / host_value, synchronising the
 * stream; with both NULL nothing is copied and the stream is not synchronised (several picks can
 * then share one read-back). */
int polar_rerank_pick(int32_t obj, int32_t dtype, const void* U, int64_t u_sj, int64_t u_si,
                      int32_t p, int32_t Br, const int32_t* sa, const int32_t* sb, const double* hh,
                      const double* om, double* g, int32_t* host_winner, double* host_value,
                      void* stream);

/* Maintenance (polar_gpu.py:243-249), applied to the caches after choosing candidate e with
 * endpoint rows (aa, bb) in the active set and solution column u (element (j,i) at
 * u + j*u_sj + i*u_si):
 *   prep:  xi[j,:] = X[j,aa]-X[j,bb]; eta[j,:] = gc - (Y[j,aa]-Y[j,bb]); D[j] = 1+u[j,aa]-u[j,bb];
 *          nxi[j] = |xi_j|^2; neta[j] = |eta_j|^2   (coef: double[p*(2q+3)] laid out xi|eta|D|nxi|neta)
 *   rank1: wX[i,j] = X[j,i,:].xi_j then X[j,i,:] -= u[j,i] xi_j / D_j   (one pass, old rows)
 *          wY[i,j] = Y[j,i,:].eta_j then Y[j,i,:] += u[j,i] eta_j / D_j;  ut[i,j] = u[j,i]
 *          (wX, wY, ut: double[nA][p], vertex-major so a pair's endpoint values are contiguous)
 *   cache: cc = (ut[ea,j]-ut[eb,j])/D_j; h = max(0, h - 2cc(wX[ea,j]-wX[eb,j]) + cc^2 nxi);
 *          t = max(0, t + 2cc(wY[ea,j]-wY[eb,j]) + cc^2 neta), one thread per pair. */
int polar_maint_prep(int32_t dtype, const void* X, const void* Y, int64_t sj, int64_t si,
                     int32_t p, int64_t q, const void* u, int64_t u_sj, int64_t u_si, int32_t aa,
                     int32_t bb, const double* gc, double* coef, void* stream);
int polar_maint_rank1(int32_t dtype, void* X, void* Y, int64_t sj, int64_t si, int32_t p,
                      int64_t nA, int64_t q, const void* u, int64_t u_sj, int64_t u_si,
                      const double* coef, double* wX, double* wY, double* ut, void* stream);
int polar_maint_cache(int32_t p, int64_t N, const int32_t* ea, const int32_t* eb, const double* coef,
                      int64_t q, const double* wX, const double* wY, const double* ut, double* h,
                      double* t, void* stream);

/* F[a, :] += g; F[b, :] -= g for one row-vector g (double[q]); F row-major, leading dim ldf. */
int polar_rows_axpy(int32_t dtype, void* F, int64_t ldf, int64_t q, int32_t a, int32_t b,
                    const double* g, void* stream);

/* Counter-hash Rademacher rows: out[i, c] = scale * (+-1), sign = top bit of
 * splitmix64(splitmix64(seed*4 + stream) + key_i*q + c), key_i = keys ? keys[i] : i. */
int polar_rademacher(int32_t dtype, uint64_t seed, int32_t stream_id, const int64_t* keys,
                     int64_t rows, int64_t q, double scale, void* out, int64_t ld, void* stream);

/* F = C0^T U for the Rademacher edge sketch, from the adjacency of op: for row i and neighbour j
 * with weight w, F[i,:] += sign * sqrt(w) * U(min(i,j)*n + max(i,j)), sign = +1 if i < j else -1. */
int polar_edge_sketch(const polar_op* op, uint64_t seed, int64_t q, double scale, void* F,
                      int64_t ldf, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* POLAR_AMD_H */
