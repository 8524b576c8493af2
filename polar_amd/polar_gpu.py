"""MI355X drop-in for realworld/polar_gpu.py (POLAR's pole-shared Krylov core).

Same entry points and argument meaning as the reference (citations per function); the numerical
work runs in the hand-written gfx950 kernels of libpolar_amd.so, called through ctypes on raw
device pointers of PyTorch-ROCm tensors.  There is no CPU fallback: without the library or a GPU
every compute entry point raises.

Additive extensions (the reference signatures stay valid):
  * `precision='fp64'|'fp32'`: dtype of the operator, vectors, shifted blocks and X/Y caches;
    scalar recurrences, inner products, h/t caches and scores are FP64 in both.
  * `poles=(tau, om)`: pole override (configs prescribe 16/24/32 poles; see quadrature.fixed_rule).
  * `width=`: columns per bank block solve (numerically neutral: columns are independent).
  * `freeze=False`: no pole freezing (the single-RHS variant of synthetic/scalability/benchmark.py).
  * results carry `marginals` (per-round G_q of the winner) and `trace_estimate`.
"""
import math
import time

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib as L
from .quadrature import composite, fixed_rule, poles  # noqa: F401  (re-exported API)
from .spmm import csr_spmm, spmm  # noqa: F401  (re-exported API; Op's product is _apply below)

__all__ = ['composite', 'fixed_rule', 'poles', 'Graph', 'Op', 'Counter', 'msolve', 'sketches',
           'chunk_width', 'polar', 'degree_greedy', 'random_set', 'true_marginals', 'spmm', 'csr_spmm',
           'RademacherRows', 'DeviceSketch']


def _device():
    L.require_gpu()
    return torch.device('cuda', torch.cuda.current_device())


def _dt(precision):
    if precision in ('fp64', torch.float64, 'float64'):
        return torch.float64
    if precision in ('fp32', torch.float32, 'float32', 'mixed'):
        return torch.float32
    raise ValueError(f'precision {precision!r}')


# ------------------------------------------------------------------------------------------------
class LaplacianCSR:
    """M0 = diag(deg + sigma) - A (polar_gpu.py:52-55).  The solver uses the graph's value-free
    adjacency + diagonal; the attributes of the reference's cupyx csr_matrix are also here:
    `.indptr`, `.indices` (int32) and `.data` (float64) are device tensors of the full CSR with the
    diagonal stored (built on first use), `.shape`, `.nnz`, `.dtype`, `.get()` (host SciPy CSR)
    and `A @ X` (the Laplacian operator kernel)."""

    def __init__(self, G):
        self._G = G
        self.shape = (G.n, G.n)
        self.dtype = np.dtype(np.float64)
        self._csr = None

    def get(self):
        G = self._G
        A = sp.csr_matrix((G._w, G._indices, G._indptr), shape=(G.n, G.n))
        M = (sp.diags(G.deg + G.sigma) - A).tocsr()
        M.sort_indices()
        return M

    def _device_csr(self):
        if self._csr is None:
            M = self.get()
            dev = _device()
            self._csr = (torch.from_numpy(M.indptr.astype(np.int32)).to(dev),
                         torch.from_numpy(M.indices.astype(np.int32)).to(dev),
                         torch.from_numpy(M.data.astype(np.float64)).to(dev))
        return self._csr

    @property
    def indptr(self):
        return self._device_csr()[0]

    @property
    def indices(self):
        return self._device_csr()[1]

    @property
    def data(self):
        return self._device_csr()[2]

    @property
    def nnz(self):
        return self._G.n + self._G.nnz

    def __matmul__(self, X):
        return spmm(self, X)


class Graph:
    """Graph(n, edges, sigma, weights=None) of polar_gpu.py:47-60, stored on the GPU."""

    def __init__(self, n, edges, sigma, weights=None):
        self.n = int(n)
        self.sigma = float(sigma)
        edges = np.asarray(edges, dtype=np.int64)
        u, v = edges[:, 0], edges[:, 1]
        w = np.ones(len(u)) if weights is None else np.asarray(weights, float)
        deg = np.bincount(u, w, self.n) + np.bincount(v, w, self.n)
        A = sp.coo_matrix((np.r_[w, w], (np.r_[u, v], np.r_[v, u])), shape=(self.n, self.n)).tocsr()
        A.sum_duplicates()
        A.sort_indices()
        # simple graph: no repeated (unordered) edge and no self-loop, i.e. every input edge kept
        # its own two adjacency entries
        self.simple = bool(A.nnz == 2 * len(u))
        if A.nnz >= 2 ** 31:
            raise ValueError('graphs with >= 2^31 adjacency entries need int64 indices')
        self._indptr = A.indptr.astype(np.int32)
        self._indices = A.indices.astype(np.int32)
        self._w = A.data.astype(np.float64)
        self.unit = bool(np.all(self._w == 1.0))
        self.deg = deg
        self.dmax = float(deg.max()) if self.n else 0.0
        self.m = len(u)
        self.edges = edges
        self.weights = w
        self.M0 = LaplacianCSR(self)
        self._dev = {}
        self._sched = None

    def Lam(self, k):
        return self.sigma + 2 * self.dmax + 4 * k

    @property
    def nnz(self):
        return int(self._indptr[-1])

    def run_rows(self):
        """Rows per work unit of the operator kernel: runs of 16 consecutive rows on graphs of 2^20
        rows and more whose mean degree is at most 8 (lattices, road networks), where a row's few
        gathers would otherwise half-fill a batch and each row pays its own index round; one row per
        unit otherwise (the denser and the smaller graphs are faster so).  Results are the same bits
        either way."""
        return 16 if self.n >= (1 << 20) and self.nnz <= 8 * self.n else 0

    def schedule(self):
        """Long-row split (rows with more than POLAR_CHUNK neighbours), host arrays."""
        if self._sched is None:
            d = np.diff(self._indptr).astype(np.int64)
            long_rows = np.flatnonzero(d > L.CHUNK).astype(np.int32)
            nch = (d[long_rows] + L.CHUNK - 1) // L.CHUNK
            base = np.zeros(len(long_rows), dtype=np.int64)
            if len(long_rows):
                base[1:] = np.cumsum(nch)[:-1]
            li = np.repeat(np.arange(len(long_rows)), nch)
            ci = np.concatenate([np.arange(c) for c in nch]) if len(nch) else np.zeros(0, np.int64)
            extra = np.stack([li, ci], axis=1).astype(np.int32) if len(li) else np.zeros((0, 2), np.int32)
            self._sched = dict(long_rows=long_rows, long_base=base.astype(np.int32),
                               extra=np.ascontiguousarray(extra), n_partial=int(nch.sum()))
        return self._sched

    def device(self, dtype):
        """Device arrays for operator dtype (indices shared between dtypes)."""
        dev = _device()
        if 'idx' not in self._dev:
            sc = self.schedule()
            self._dev['idx'] = dict(
                indptr=torch.from_numpy(self._indptr).to(dev),
                indices=torch.from_numpy(self._indices).to(dev) if self.nnz else torch.zeros(1, dtype=torch.int32, device=dev),
                long_rows=torch.from_numpy(sc['long_rows']).to(dev),
                long_base=torch.from_numpy(sc['long_base']).to(dev),
                extra=torch.from_numpy(sc['extra'].reshape(-1)).to(dev),
            )
        if dtype not in self._dev:
            self._dev[dtype] = dict(
                diag=torch.from_numpy(self.deg + self.sigma).to(dev, dtype),
                weights=None if self.unit else torch.from_numpy(self._w).to(dev, dtype),
            )
        return self._dev['idx'], self._dev[dtype]


class Op:
    """y = (M0 + sum_S b b^T + shift I) x with unit-weight added edges (polar_gpu.py:63-80).
    The added edges form a row patch applied inside the SpMM kernel (deterministic; no atomics)."""

    def __init__(self, G, precision='fp64'):
        self.G = G
        self.M0 = G.M0
        self.A = []
        self.B = []
        self.grp = []             # column group of each added edge (None: all columns)
        self.dtype = _dt(precision)
        self._patch = None
        self._struct = None
        self.cpl_max = 0          # launch-shape tuning (polar_amd.autotune); 0 = library default
        self.grid_max = 0
        self.run_rows = 0         # rows per operator work unit; 0 = Graph.run_rows()
        self.use_graph = False

    def add(self, a, b, group=None):
        """Add the unit-weight edge (a, b) (polar_gpu.py:68-69); with `group`, only for the columns of
        that column group (several systems on one graph, see polar_amd.batched.polar_columns)."""
        self.A.append(int(a))
        self.B.append(int(b))
        self.grp.append(None if group is None else int(group))
        self._patch = None
        self._struct = None

    def _patch_arrays(self):
        if self._patch is None:
            dev = _device()
            if self.A:
                nbr = {}
                for a, b, g in zip(self.A, self.B, self.grp):
                    nbr.setdefault(a, []).append((b, g))
                    nbr.setdefault(b, []).append((a, g))
                rows = np.array(sorted(nbr), dtype=np.int32)
                ptr = np.zeros(len(rows) + 1, dtype=np.int32)
                ptr[1:] = np.cumsum([len(nbr[r]) for r in rows])
                cols = np.array([c for r in rows for c, _ in nbr[r]], dtype=np.int32)
                parts = [rows, ptr, cols]
                if any(g is not None for g in self.grp):
                    parts.append(np.array([-1 if g is None else g for r in rows for _, g in nbr[r]], dtype=np.int32))
                # one upload for all the patch arrays (views of one device buffer): each pageable
                # copy waits for the stream, and the greedy rebuilds the patch every round
                buf = torch.from_numpy(np.concatenate(parts)).to(dev)
                offs = np.cumsum([0] + [len(x) for x in parts])
                self._patch = tuple(buf[offs[i]:offs[i + 1]] for i in range(len(parts)))
            else:
                self._patch = ()
        return self._patch

    def c_op(self, dtype=None):
        """ctypes polar_op for this operator (kept alive by self)."""
        dtype = self.dtype if dtype is None else dtype
        if self._struct is not None and self._struct[0] == dtype:
            return self._struct[1]
        idx, dd = self.G.device(dtype)
        sc = self.G.schedule()
        patch = self._patch_arrays()
        s = L.PolarOp()
        s.n = self.G.n
        s.dtype = L.dtype_code(dtype)
        s.weighted = 0 if dd['weights'] is None else 1
        s.indptr = L.ptr(idx['indptr'])
        s.indices = L.ptr(idx['indices'])
        s.weights = L.ptr(dd['weights'])
        s.diag = L.ptr(dd['diag'])
        s.n_long = len(sc['long_rows'])
        s.n_extra = len(sc['extra'])
        s.n_partial = sc['n_partial']
        s.chunk = L.CHUNK
        s.cpl_max = self.cpl_max
        s.grid_max = self.grid_max
        s.run_rows = self.run_rows if self.run_rows else self.G.run_rows()
        s.long_rows = L.ptr(idx['long_rows']) if s.n_long else None
        s.long_base = L.ptr(idx['long_base']) if s.n_long else None
        s.extra = L.ptr(idx['extra']) if s.n_extra else None
        if patch:
            s.n_patch = len(patch[0])
            s.patch_rows, s.patch_ptr, s.patch_cols = (L.ptr(t) for t in patch[:3])
            s.patch_grp = L.ptr(patch[3]) if len(patch) > 3 else None
        self._struct = (dtype, s)
        return s

    def __call__(self, x, shift):
        return _apply(self, x, shift)


def _apply(op, x, shift):
    lib = L.require_gpu()
    x = torch.as_tensor(x, device=_device())
    one = x.ndim == 1
    X = x.reshape(op.G.n, -1).to(op.dtype).contiguous()
    R = X.shape[1]
    Y = torch.empty_like(X)
    s = op.c_op(X.dtype)
    ws = torch.empty(max(lib.polar_spmm_ws_bytes(s, min(R, 64)), 256), dtype=torch.uint8, device=X.device)
    for lo in range(0, R, 64):
        r = min(64, R - lo)
        L.check(lib.polar_spmm(s, float(shift or 0.0), L.ptr(X[:, lo:]), R, L.ptr(Y[:, lo:]), R, r,
                               L.ptr(ws), ws.numel(), L.stream_ptr()), 'polar_spmm')
    return Y.reshape(-1) if one else Y


class Counter:
    """polar_gpu.py:83-86 (plus `solves`: column-iterations of each msolve call)."""

    def __init__(self):
        self.spmv = 0
        self.rhs = 0
        self.solves = []


def _max_cols(dtype):
    return 128 if dtype == torch.float64 else 256


def _blocked(block, dtype):
    """Blocked shifted update: on by default for FP32, off for FP64 (the pinned path)."""
    return (dtype == torch.float32) if block is None else bool(block)


def _replacing(replace, block, dtype):
    """FP64 residual replacement in the greedy's solves: on by default for the FP32-mixed form (with
    its blocked update, where the seed iterate advances once per window), off for FP64."""
    if replace is None:
        return dtype == torch.float32 and _blocked(block, dtype)
    return bool(replace)


def _col_bytes(n, p, nA, dtype, block=None, replace=False):
    """Device bytes per solved column: the n-row CG vectors (with the residual ring of the blocked
    update), the FP64 seed iterate, FP64 product and saved direction of residual replacement, and
    the shifted P block."""
    esz = 8 if dtype == torch.float64 else 4
    nvec = 11 if _blocked(block, dtype) else 4
    return esz * (nvec * n + p * nA) + ((16 + esz) * n if replace else 0) + 64


HISTORY_BYTES = 64 << 30      # residual history budget of one re-rank solve ...
HISTORY_SHARE = 0.25          # ... and at most this share of the device memory still available
HISTORY_STEPS = 1024          # iterations a history solve keeps at most


def _available_bytes():
    """Device memory free for new tensors: the driver's free memory plus what torch's caching
    allocator holds unused."""
    free, _ = torch.cuda.mem_get_info()
    return free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated()


def history_cap(n, R, dtype, block=None, avail=None):
    """Iterations a re-rank solve of R columns on n rows keeps in history mode: with the blocked
    update (the FP32 default, FP64 with block=True) when at least 256 iterations fit in
    min(HISTORY_BYTES, HISTORY_SHARE x available device memory), else 0 (the solve forms its
    shifted blocks as usual).  The pinned FP64 path never uses it: its per-iteration update is the
    oracle's arithmetic.  avail: available bytes (None: asked from the device)."""
    if not _blocked(block, dtype):
        return 0
    esz = 8 if dtype == torch.float64 else 4
    if avail is None:
        avail = _available_bytes()
    budget = min(HISTORY_BYTES, int(HISTORY_SHARE * avail))
    cap = min(HISTORY_STEPS, budget // max(1, n * R * esz))
    return int(cap) if cap >= 256 else 0


class History:
    """A solve run with POLAR_HISTORY: every residual and every pole's per-step coefficients are kept
    in its workspace, and X_j is formed only where it is needed -- on two rows of every column (the
    re-rank's endpoint values, polar_gpu.py:230) and on every active row of the winner's column
    (:237), instead of on every active row of every column during the solve."""

    def __init__(self, op, s, R, p, nA, flags, tcap, ws, dtype):
        self.op, self.s, self.R, self.p, self.nA = op, s, R, p, nA
        self.flags, self.tcap, self.ws, self.dtype = flags, tcap, ws, dtype

    def _args(self):
        return (self.s, self.R, self.p, self.nA, self.flags, self.tcap, L.ptr(self.ws), self.ws.numel())

    def rows(self, rows):
        """E[j, k, c] = X_j[rows[k, c], c]; rows: device int32 (K, R) absolute row ids."""
        lib = L.require_gpu()
        rows = rows.to(torch.int32).contiguous()
        K = rows.shape[0]
        E = torch.empty((self.p, K, self.R), dtype=self.dtype, device=self.ws.device)
        L.check(lib.polar_history_rows(*self._args(), K, L.ptr(rows), L.ptr(E), L.stream_ptr()), 'polar_history_rows')
        return E

    def cols(self, cols, active):
        """X[i, j, a] = X_j[active[a], cols[i]]; cols: device int32 (C,), active: device int32 (nA,)."""
        lib = L.require_gpu()
        cols = cols.to(torch.int32).contiguous()
        X = torch.empty((cols.numel(), self.p, self.nA), dtype=self.dtype, device=self.ws.device)
        for lo in range(0, cols.numel(), 4):        # at most 4 columns per call
            C = min(4, cols.numel() - lo)
            L.check(lib.polar_history_cols(*self._args(), C, L.ptr(cols[lo:]), L.ptr(active), L.ptr(X[lo:]),
                                           L.stream_ptr()), 'polar_history_cols')
        return X


def msolve(op, W, tau, active, tol, maxit=100000, track=False, counter=None, freeze=True,
           trace=False, out=None, width=None, graph=None, block=None, replace=False, history=0):
    """Block multi-shift CG for (M + tau_j I) U = W on all poles (polar_gpu.py:88-133).

    Returns (X[p, nA, R], h[p, R] or None, iters[R] host int64) and, with trace=True, a 4th value
    te[p, R] = w^T (M + tau_j)^{-1} w.  `out` may be a (p, nA, R) view with unit column stride
    into which X is written (the bank).  `width` caps the columns solved together (the workspace
    scales with it); columns are independent, so the results do not depend on it.  graph=True
    replays the CG iterations as hipGraphs of 8 iterations (same results, fewer launch gaps);
    None takes the operator's tuned choice (op.use_graph, see polar_amd.autotune).  block=True
    applies the shifted x/p updates once per 8 iterations from the stored residuals (same CG
    scalars, counts and tracked norms; X within rounding; ~8x fewer passes over the shifted
    blocks); None = on for FP32, off for FP64 (whose default path is the pinned one).
    replace=True: residual replacement every 24 iterations (r = w - (M + tau0) x0 with the seed
    iterate x0 accumulated in FP64 and the product in FP64).  history=T > 0: POLAR_HISTORY with T
    iterations at most (maxit is then T; more raises the iteration-cap error): X is not formed, the
    first value returned is None and a History is appended to the results."""
    lib = L.require_gpu()
    dev = _device()
    dtype = op.dtype
    W = torch.as_tensor(W, device=dev)
    if W.ndim == 1:
        W = W[:, None]
    W = W.to(dtype)
    if W.stride(1) != 1:
        W = W.contiguous()
    n, R = W.shape
    tau_h = np.asarray(tau, dtype=np.float64)
    p = len(tau_h)
    tau_d = torch.from_numpy(tau_h).to(dev)
    if torch.is_tensor(active) and active.is_cuda and active.dtype == torch.int32:
        act_d = active.contiguous()
    else:
        act_d = torch.as_tensor(np.asarray(active.cpu() if torch.is_tensor(active) else active),
                                dtype=torch.int32).to(dev)
    nA = act_d.numel()
    if history:
        X = None
    else:
        X = out if out is not None else torch.empty((p, nA, R), dtype=dtype, device=dev)
        assert X.shape == (p, nA, R) and X.stride(2) == 1
    h = torch.empty((p, R), dtype=torch.float64, device=dev) if track else None
    te = torch.empty((p, R), dtype=torch.float64, device=dev) if trace else None
    iters = torch.empty(R, dtype=torch.int64, device=dev)
    if graph is None:
        graph = getattr(op, 'use_graph', False)
    flags = ((L.POLAR_TRACK if track else 0) | (L.POLAR_TRACE if trace else 0) |
             (0 if freeze else L.POLAR_NOFREEZE) | (L.POLAR_GRAPH if graph else 0) |
             (L.POLAR_BLOCK8 if _blocked(block, dtype) else 0) | (L.POLAR_REPLACE if replace else 0))
    s = op.c_op(dtype)
    step = min(R, _max_cols(dtype), width or R)
    if history:
        assert R <= _max_cols(dtype)
        step, maxit = R, int(history)
        flags |= L.POLAR_HISTORY
        ws_bytes = lib.polar_msolve_history_ws_bytes(s, R, p, nA, flags, maxit)
    else:
        ws_bytes = lib.polar_msolve_ws_bytes(s, step, p, nA, flags)
    L.check(0 if ws_bytes > 0 else L.POLAR_E_ARG, 'msolve workspace size')
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    for lo in range(0, R, step):
        r = min(step, R - lo)
        # h/te blocks are written as contiguous p x r arrays
        hb = (h if r == R else torch.empty((p, r), dtype=torch.float64, device=dev)) if track else None
        tb = (te if r == R else torch.empty((p, r), dtype=torch.float64, device=dev)) if trace else None
        rc = lib.polar_msolve(s, L.ptr(W[:, lo:]), W.stride(0), r, L.ptr(tau_d), p, L.ptr(act_d), nA,
                              float(tol), int(maxit), flags, None if X is None else L.ptr(X[:, :, lo:]),
                              0 if X is None else X.stride(0), 0 if X is None else X.stride(1), L.ptr(hb),
                              L.ptr(tb), L.ptr(iters[lo:]), L.ptr(ws), ws_bytes, L.stream_ptr())
        L.check(rc, 'polar_msolve')
        if track and r != R:
            h[:, lo:lo + r] = hb
        if trace and r != R:
            te[:, lo:lo + r] = tb
    it_h = iters.cpu().numpy()
    if counter is not None:
        counter.spmv += int(it_h.sum())
        counter.rhs += R
        counter.solves.append(int(it_h.sum()))
    res = (X, h, it_h)
    if trace:
        res = res + (te,)
    if history:
        res = res + (History(op, s, R, p, nA, flags, maxit, ws, dtype),)
    return res


# ------------------------------------------------------------------------------------------------
def _block_rows(rows, q, rng, center=False):
    """Stacked orthogonal q x q row blocks (polar_gpu.py:137-145), host NumPy."""
    nb = -(-rows // q)
    Wb = rng.normal(size=(nb, q, q))
    Q, Rm = np.linalg.qr(Wb)
    Q = Q * np.sign(np.diagonal(Rm, axis1=1, axis2=2))[:, None, :]
    Wb = np.transpose(Q, (0, 2, 1)).reshape(nb * q, q)[:rows]
    if center:
        Wb = Wb - Wb.mean(0)
    return Wb


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return x ^ (x >> np.uint64(31))


def _rad_rows(seed, stream, keys, q):
    with np.errstate(over='ignore'):
        base = _splitmix64(np.uint64(seed * 4 + stream))
        k = np.asarray(keys, dtype=np.uint64)[:, None] * np.uint64(q) + np.arange(q, dtype=np.uint64)[None]
        hsh = _splitmix64((base + k) & _M64)
    return np.where((hsh >> np.uint64(63)) == 1, -1.0, 1.0)


class RademacherRows:
    """Lazily generated candidate sketch rows Gc[e] (counter hash, stream 2)."""

    def __init__(self, seed, q, N):
        self.seed, self.q, self.N = int(seed), int(q), int(N)

    def __len__(self):
        return self.N

    def __getitem__(self, e):
        return _rad_rows(self.seed, 2, np.atleast_1d(e), self.q)[0] / math.sqrt(self.q)

    def full(self):
        return _rad_rows(self.seed, 2, np.arange(self.N), self.q) / math.sqrt(self.q)

    def fill_device(self, keys, out):
        """out[i] = Gc[keys[i]] on the GPU (FP64; keys: device int64 candidate indices, out: device
        (len(keys), q) with unit column stride), the same values as __getitem__."""
        lib = L.require_gpu()
        L.check(lib.polar_rademacher(L.POLAR_F64, self.seed, 2, L.ptr(keys), len(keys), self.q,
                                     1.0 / math.sqrt(self.q), L.ptr(out), out.stride(0), L.stream_ptr()),
                'polar_rademacher')


class DeviceSketch:
    """A counter-hash Rademacher sketch (Z, F, Gc) that is generated on the GPU where polar() needs
    it, so no resident copy is held beside the solver's own right-hand-side blocks."""

    def __init__(self, G, seed, q, N):
        self.G, self.seed, self.q = G, int(seed), int(q)
        self.rows = RademacherRows(seed, q, N)

    def fill(self, Zout, Fout, op):
        """Write Z and F (n x q views with unit column stride) in op's dtype."""
        lib = L.require_gpu()
        scale = 1.0 / math.sqrt(self.q)
        dt = L.dtype_code(Zout.dtype)
        L.check(lib.polar_rademacher(dt, self.seed, 0, None, self.G.n, self.q, scale, L.ptr(Zout), Zout.stride(0),
                                     L.stream_ptr()), 'polar_rademacher')
        L.check(lib.polar_edge_sketch(op.c_op(Fout.dtype), self.seed, self.q, scale, L.ptr(Fout), Fout.stride(0),
                                      L.stream_ptr()), 'polar_edge_sketch')


def sketches(G, cand, q, seed, kind='block', chunk=4_000_000, on_device=False, precision='fp64'):
    """(Z n x q, F = C0^T U n x q, Gc N x q) of polar_gpu.py:148-165.  kind 'block' / 'gauss' follow
    the reference on the host; kind 'rademacher' uses the counter hash (+-1/sqrt(q)), which the
    GPU can generate in place (on_device=True: Z and F are device tensors, Gc rows are generated
    per use; on_device='lazy': a DeviceSketch that polar() generates into its own blocks)."""
    n = G.n
    if kind == 'rademacher':
        scale = 1.0 / math.sqrt(q)
        if on_device and not G.simple:
            # the host form adds one hashed row per input edge; the device kernel hashes each merged
            # adjacency entry once, which differs for a repeated edge or a self-loop
            raise ValueError('the device Rademacher sketch needs a simple graph (no repeated edge, no '
                             'self-loop); use on_device=False for a multigraph')
        if on_device == 'lazy':
            return DeviceSketch(G, seed, q, len(cand))
        if on_device:
            lib = L.require_gpu()
            dev = _device()
            dtype = _dt(precision)
            Z = torch.empty((n, q), dtype=dtype, device=dev)
            F = torch.empty((n, q), dtype=dtype, device=dev)
            L.check(lib.polar_rademacher(L.dtype_code(dtype), int(seed), 0, None, n, q, scale, L.ptr(Z), q,
                                         L.stream_ptr()), 'polar_rademacher')
            op = Op(G, precision=dtype)
            L.check(lib.polar_edge_sketch(op.c_op(dtype), int(seed), q, scale, L.ptr(F), q, L.stream_ptr()),
                    'polar_edge_sketch')
            return Z, F, RademacherRows(seed, q, len(cand))
        Z = _rad_rows(seed, 0, np.arange(n), q) * scale
        F = np.zeros((n, q))
        u = np.minimum(G.edges[:, 0], G.edges[:, 1])
        v = np.maximum(G.edges[:, 0], G.edges[:, 1])
        sw = np.sqrt(G.weights)
        for lo in range(0, G.m, chunk):
            hi = min(G.m, lo + chunk)
            U = _rad_rows(seed, 1, u[lo:hi] * n + v[lo:hi], q) * (scale * sw[lo:hi, None])
            np.add.at(F, u[lo:hi], U)
            np.add.at(F, v[lo:hi], -U)
        return Z, F, RademacherRows(seed, q, len(cand)).full()
    rng = np.random.default_rng(seed)
    if kind == 'gauss':
        Z = rng.normal(size=(n, q)) / np.sqrt(q)
    else:
        Z = _block_rows(n, q, rng, center=True)
    F = np.zeros((n, q))
    u, v = G.edges[:, 0], G.edges[:, 1]
    sw = np.sqrt(G.weights)
    for lo in range(0, G.m, chunk):
        hi = min(G.m, lo + chunk)
        U = rng.normal(size=(hi - lo, q)) / np.sqrt(q) if kind == 'gauss' else _block_rows(hi - lo, q, rng)
        U *= sw[lo:hi, None]
        C = sp.coo_matrix((np.r_[np.ones(hi - lo), -np.ones(hi - lo)],
                           (np.r_[u[lo:hi], v[lo:hi]], np.r_[np.arange(hi - lo), np.arange(hi - lo)])),
                          shape=(n, hi - lo)).tocsr()
        F += C @ U
    Gc = rng.normal(size=(len(cand), q)) / np.sqrt(q) if kind == 'gauss' else _block_rows(len(cand), q, rng)
    return Z, F, Gc


def chunk_width(n, p, nA, budget=10e9):
    """polar_gpu.py:169-173 (the reference's column budget; kept for API parity)."""
    per_col = 8.0 * (10 * n + 4 * p * nA)
    cap = 16 if n > 3e6 else 64
    return int(max(8, min(cap, budget // per_col)))


def bank_width(n, p, nA, q, dtype, budget=None, block=None, replace=False):
    """Columns per bank block solve on one MI355X: as wide as one wave covers (128 FP64 / 256
    FP32 columns) within a memory budget for the shifted P block and the CG vectors."""
    return fit_width(n, p, nA, 2 * q, dtype, budget, block, replace)


def fit_width(n, p, nA, R, dtype, budget=None, block=None, replace=False):
    """Columns (at most R) whose msolve workspace fits in `budget` bytes (default: 70% of the
    free device memory): per column the n-row CG vectors (with the residual ring of the blocked
    update and the vectors of residual replacement) and the shifted P block."""
    esz = 8 if dtype == torch.float64 else 4
    if budget is None:
        free, _ = torch.cuda.mem_get_info()
        budget = 0.7 * free
    per_col = _col_bytes(n, p, nA, dtype, block, replace)
    w = int(min(R, _max_cols(dtype), max(1, budget // per_col)))
    # equal chunks (e.g. 2 x 64, not 120 + 8), multiples of 8 where possible
    chunks = -(-R // w)
    w = -(-R // chunks)
    if w > 8 and w % 8 and (w + 7) // 8 * 8 * chunks - R < 8 * chunks and (w + 7) // 8 * 8 * per_col <= budget:
        w = (w + 7) // 8 * 8
    # When the block is split, each chunk writes a column slice of a wider (p, nA, R) block (the
    # bank's X/Y caches): a width of whole 128-byte lines keeps every slice line-aligned, so the
    # per-iteration shifted updates move no partial lines (c5 FP64 bank: 22 columns 146 s, 16
    # columns 106 s).  Taken when it keeps at least 60% of the width.
    lc = 128 // esz
    if chunks > 1 and w > lc and w % lc and (w // lc) * lc >= 0.6 * w:
        w = (w // lc) * lc
    return max(1, min(w, R))


def _sync():
    torch.cuda.synchronize()


class _Phases:
    """Optional per-phase wall-clock breakdown of polar() (device-synchronised marks)."""

    def __init__(self, on):
        self.on = on
        self.acc = {}
        self.t = time.perf_counter()

    def mark(self, name):
        if not self.on:
            return
        torch.cuda.synchronize()
        now = time.perf_counter()
        self.acc[name] = self.acc.get(name, 0.0) + (now - self.t)
        self.t = now


# ------------------------------------------------------------------------------------------------
def polar(G, cand, k, s, q=128, B=16, eps=1e-2, tol=1e-7, obj='frac', sketch=None, naive=False,
          seed=2026, kind='block', trace=None, time_limit=None, poles=None, width=None,
          freeze=True, precision='fp64', profile=False, progress=None, graph=None, tune=None, block=None,
          shortlist='proxy', replace=None, history=None):
    """POLAR greedy, cache backend (polar_gpu.py:180-254), on the GPU.

    profile=True adds `phases`: seconds per phase (bank, select, rerank_solve, rerank_pick,
    maintenance, update), measured with a device synchronisation at each phase boundary.
    graph=True replays the CG iterations of every solve as hipGraphs (same results).
    tune: a launch-shape dict from polar_amd.autotune (same results) or True to tune here (before
    the timed region).  block: blocked shifted update in every solve (see msolve; None = FP32 on,
    FP64 off).  shortlist='gradient' ranks candidates by the first-order gain sum_j om_j h[j,e]
    (minus the sketched trace's derivative, see polar_amd.adjoint) instead of the reference's
    proxy; the re-rank by exact marginals is unchanged.  replace: FP64 residual replacement every
    24 iterations in every solve (see msolve); None = on for FP32 (the FP32-mixed form), off for FP64.
    history: the re-rank solve with the blocked update keeps its residuals and forms the shifted
    solutions only on the shortlist's endpoint rows and the winner's column (msolve history=,
    History); None = when its residual history fits (history_cap), False = never (the re-rank forms
    every column's shifted blocks, as the pinned FP64 path and polar_amd.batched.polar_batched
    always do).  Both agree within the rounding of the vectors' precision.
    trace: a list that receives per round dict(shortlist, chosen) as in the reference, plus `proxy`
    (the shortlist's proxy scores) and `scores` (its re-rank scores; empty for naive).
    progress: optional callable(stage, info) called after the bank ('bank', {seconds, spmv}) and
    after each round ('round', {round, edge, marginal, seconds, spmv}); it reads host values only.

    Returns dict(sequence, seconds, spmv, rhs, p, rounds, truncated, marginals, trace_estimate,
    solves, precision, proxy_gap, rerank_gap, replace).  proxy_gap[r] / rerank_gap[r]: the relative
    gap (top1 - top2) / |top1| between the two best proxy scores and between the two best re-rank
    scores of round r (the reference's picks are the first of a stable argsort, :221, and the
    first argmax, :236, so a gap at the CG-tolerance level marks a near-tie); NaN where fewer than
    two candidates (or shortlisted candidates) compete."""
    lib = L.require_gpu()
    dev = _device()
    dtype = _dt(precision)
    replace = _replacing(replace, block, dtype)
    dcode = L.dtype_code(dtype)
    stream = L.stream_ptr()
    n, N = G.n, len(cand)
    sigma = G.sigma
    if poles is None:
        tau, om, _ = globals()['poles'](obj, s, sigma, G.Lam(k), eps)
    else:
        tau, om = np.asarray(poles[0], float), np.asarray(poles[1], float)
    p = len(tau)
    if sketch is None:
        sketch = sketches(G, cand, q, seed, kind)
    lazy = isinstance(sketch, DeviceSketch)
    if lazy:
        Z = F = None
        Gc = sketch.rows
        q = sketch.q
    else:
        Z, F, Gc = sketch
        q = Z.shape[1]
    cand = np.asarray(cand)
    active_h = np.unique(cand)                                                    # :191
    es_h = np.searchsorted(active_h, cand)
    nA = len(active_h)
    active = torch.from_numpy(active_h.astype(np.int32)).to(dev)
    ea = torch.from_numpy(es_h[:, 0].astype(np.int32)).to(dev)
    eb = torch.from_numpy(es_h[:, 1].astype(np.int32)).to(dev)
    tau_d = torch.from_numpy(tau).to(dev)
    om_d = torch.from_numpy(om).to(dev)
    # Bank right-hand sides [Z | F]: one n x 2q block when one block solve covers all 2q columns
    # with room to spare, else separate Z and F blocks (Z is released after the bank).  Columns
    # are independent, so the split does not change any result.
    esz = 8 if dtype == torch.float64 else 4
    free = torch.cuda.mem_get_info()[0]
    joint = (p * nA * 2 * q + 2 * n * q) * esz + 2 * q * _col_bytes(n, p, nA, dtype, block, replace) < 0.8 * free
    if joint:
        RHS = torch.empty((n, 2 * q), dtype=dtype, device=dev)
        Zb, Fb = RHS[:, :q], RHS[:, q:]
    else:
        Zb = torch.empty((n, q), dtype=dtype, device=dev)
        Fb = torch.empty((n, q), dtype=dtype, device=dev)
    op = Op(G, precision=dtype)
    if lazy:
        sketch.fill(Zb, Fb, op)
    else:
        Zb.copy_(torch.as_tensor(Z, device=dev).to(dtype))
        Fb.copy_(torch.as_tensor(F, device=dev).to(dtype))
    Zbox = [Zb]
    del Zb
    Gc_dev = None
    if isinstance(Gc, torch.Tensor):
        Gc_dev = Gc.to(dev, torch.float64)
    elif not isinstance(Gc, RademacherRows):
        Gc = np.asarray(Gc, dtype=np.float64)
    if tune is True:
        from . import autotune
        tune = autotune.tune(op, min(2 * q, _max_cols(dtype)), tau, active)
    if tune:
        from . import autotune
        autotune.apply(op, tune)
    cnt = Counter()
    live = np.ones(N, bool)
    dead = torch.zeros(N, dtype=torch.uint8, device=dev)
    seq, marg = [], []
    traces = []
    want_trace = obj != 'logdet'
    objc = 2 if obj == 'logdet' else 0
    score = torch.empty(N, dtype=torch.float64, device=dev)
    sel_ws = torch.empty(lib.polar_select_ws_bytes(N), dtype=torch.uint8, device=dev)
    sel_out = torch.empty(64, dtype=torch.int32, device=dev)
    gaps_d = torch.full((max(k, 1), 4), float('nan'), dtype=torch.float64, device=dev)   # top-2 proxy, top-2 re-rank
    coef = torch.empty((p, 2 * q + 3), dtype=torch.float64, device=dev)
    wX = wY = wU = None                 # maintenance buffers, allocated after the bank solves
    gvec = torch.empty(max(B, 1) + 2, dtype=torch.float64, device=dev)
    # endpoint row indices 0 / 1 of a (p, 2, Br) block of endpoint values (history re-rank)
    z01 = torch.arange(2, dtype=torch.int32, device=dev)[:, None].repeat(1, max(B, 1)).contiguous()
    gc_row = torch.empty(q, dtype=torch.float64, device=dev)
    tr_hist = []
    _sync()
    t0 = time.perf_counter()                                                      # :196
    ph = _Phases(profile)

    XY = torch.empty((p, nA, 2 * q), dtype=dtype, device=dev)
    if width is None:
        width = bank_width(n, p, nA, q, dtype, block=block, replace=replace)   # sized on what is free after the caches

    if joint:
        blocks = [(lambda: RHS, 0, 2 * q)]
    else:
        blocks = [(lambda: Zbox[0], 0, q), (lambda: Fb, q, 2 * q)]

    def bank():                                                                    # :200-206
        terms = []
        chunks = [(src, c0, lo, min(c1, lo + width)) for src, c0, c1 in blocks for lo in range(c0, c1, width)]
        for src, c0, lo, hi in chunks:
            res = msolve(op, src()[:, lo - c0:hi - c0], tau, active, tol, counter=cnt, freeze=freeze,
                         trace=want_trace and lo < q, out=XY[:, :, lo:hi], graph=graph, block=block,
                         replace=replace)
            if want_trace and lo < q:
                ncol = min(q, hi) - lo
                terms.append((om[:, None] * res[3][:, :ncol].cpu().numpy()).ravel())
        tr = math.fsum(np.concatenate(terms)) if terms else 0.0   # correctly rounded
        X, Y = XY[:, :, :q], XY[:, :, q:]
        h = torch.empty((p, N), dtype=torch.float64, device=dev)
        t = torch.empty((p, N), dtype=torch.float64, device=dev)
        L.check(lib.polar_pair_sqnorm(dcode, L.ptr(X), X.stride(0), X.stride(1), p, q, L.ptr(ea), L.ptr(eb), N,
                                      L.ptr(h), stream), 'pair_sqnorm')
        L.check(lib.polar_pair_sqnorm(dcode, L.ptr(Y), Y.stride(0), Y.stride(1), p, q, L.ptr(ea), L.ptr(eb), N,
                                      L.ptr(t), stream), 'pair_sqnorm')
        return X, Y, h, t, tr

    Ubuf = None
    hist_ok = True                      # cleared once a history solve outruns its history
    X, Y, h, t, tr = bank()
    if not naive and not joint:
        Zbox[0] = None                      # Z is not needed again
    ph.mark('bank')
    if progress is not None:
        progress('bank', dict(seconds=time.perf_counter() - t0, spmv=cnt.spmv))
    traces.append(tr if want_trace else None)
    for rnd in range(k):
        if rnd and time_limit is not None and time.perf_counter() - t0 > time_limit:   # :215
            break
        if naive and rnd:
            X, Y, h, t, tr = bank()
            traces.append(tr if want_trace else None)
            ph.mark('bank')
        L.check(lib.polar_proxy(3 if shortlist == 'gradient' else objc, L.ptr(h), L.ptr(t), L.ptr(tau_d),
                                L.ptr(om_d), p, N, sigma, L.ptr(dead), L.ptr(score), stream), 'proxy')
        nlive = int(live.sum())
        Br = 1 if naive else min(B, nlive)
        Bsel = min(max(Br, 2), nlive)        # at least the top two, for the proxy gap record
        short = np.zeros(Bsel, dtype=np.int32)
        L.check(lib.polar_select_topb(L.ptr(score), N, Bsel, L.ptr(sel_out), short.ctypes.data_as(L.vp),
                                      L.ptr(sel_ws), sel_ws.numel(), stream), 'select_topb')   # :221
        short = short[:Br]
        gaps_d[rnd, :min(Bsel, 2)] = score[sel_out[:min(Bsel, 2)].long()]
        if naive:
            e = int(short[0])
            marg.append(float(score[e].item()))
        else:
            a_idx = torch.from_numpy(cand[short, 0].astype(np.int32)).to(dev)
            b_idx = torch.from_numpy(cand[short, 1].astype(np.int32)).to(dev)
            Wb = torch.empty((n, Br), dtype=dtype, device=dev)
            L.check(lib.polar_pair_rhs(dcode, n, Br, L.ptr(a_idx), L.ptr(b_idx), L.ptr(Wb), stream), 'pair_rhs')
            need_track = obj != 'logdet'
            ph.mark('select')
            H = None
            hcap = history_cap(n, Br, dtype, block) if history is not False and hist_ok else 0
            if hcap:
                # FP32: keep the residuals, form X_j only at the shortlisted endpoints and the winner
                try:
                    _, hh, _, H = msolve(op, Wb, tau, active, tol, track=need_track, counter=cnt, freeze=freeze,
                                         graph=graph, block=block, replace=replace, history=hcap)
                except RuntimeError as ex:
                    if 'iteration cap' not in str(ex):
                        raise
                    hist_ok = False             # the later re-rank solves need about as many
            if H is None:
                if Ubuf is None:                # one re-rank block for all rounds
                    Ubuf = torch.empty((p, nA, max(B, 1)), dtype=dtype, device=dev)
                U = Ubuf[:, :, :Br]
                rw = fit_width(n, p, nA, Br, dtype, block=block)
                U, hh, _ = msolve(op, Wb, tau, active, tol, track=need_track, counter=cnt, freeze=freeze, out=U,
                                  width=rw, graph=graph, block=block, replace=replace)[:3]
                sa = torch.from_numpy(es_h[short, 0].astype(np.int32)).to(dev)
                sb = torch.from_numpy(es_h[short, 1].astype(np.int32)).to(dev)
            else:
                U = H.rows(torch.stack([a_idx, b_idx]))        # (p, 2, Br): U at the two endpoints
                sa, sb = z01[0], z01[1]
            ph.mark('rerank_solve')
            win = np.zeros(1, dtype=np.int32)
            val = np.zeros(1, dtype=np.float64)
            L.check(lib.polar_rerank_pick(objc, dcode, L.ptr(U), U.stride(0), U.stride(1), p, Br, L.ptr(sa), L.ptr(sb),
                                          L.ptr(hh), L.ptr(om_d), L.ptr(gvec), win.ctypes.data_as(L.vp),
                                          val.ctypes.data_as(L.vp), stream), 'rerank_pick')
            j = 0 if Br == 1 else int(win[0])
            e = int(short[j])
            if H is None:
                u = U[:, :, j]
            elif rnd + 1 < k:                   # the winner's column, for the maintenance
                u = H.cols(torch.full((1,), j, dtype=torch.int32, device=dev), active)[0]
            marg.append(float(val[0]))
            if Br >= 2:
                gaps_d[rnd, 2:4] = torch.topk(gvec[:Br], 2).values
        if trace is not None:
            trace.append(dict(shortlist=[int(x) for x in short], chosen=e,
                              proxy=score[torch.from_numpy(short.astype(np.int64)).to(dev)].cpu().tolist(),
                              scores=[] if naive else gvec[:Br].cpu().tolist()))
        seq.append(e)
        live[e] = False
        dead[e] = 1
        aa, bb = (int(x) for x in es_h[e])
        a, b = (int(x) for x in cand[e])
        if Gc_dev is not None:
            gc_row.copy_(Gc_dev[e])
        else:
            gc_row.copy_(torch.from_numpy(np.ascontiguousarray(Gc[e], dtype=np.float64)))
        ph.mark('rerank_pick')
        if not naive and rnd + 1 < k:                                              # :242-249
            if wX is None:              # 3 p nA doubles, kept out of the bank's memory budget
                wX = torch.empty((p, nA), dtype=torch.float64, device=dev)
                wY = torch.empty((p, nA), dtype=torch.float64, device=dev)
                wU = torch.empty((nA, p), dtype=torch.float64, device=dev)
            L.check(lib.polar_maint_prep(dcode, L.ptr(X), L.ptr(Y), X.stride(0), X.stride(1), p, q, L.ptr(u),
                                         u.stride(0), u.stride(1), aa, bb, L.ptr(gc_row), L.ptr(coef), stream),
                    'maint_prep')
            L.check(lib.polar_maint_rank1(dcode, L.ptr(X), L.ptr(Y), X.stride(0), X.stride(1), p, nA, q, L.ptr(u),
                                          u.stride(0), u.stride(1), L.ptr(coef), L.ptr(wX), L.ptr(wY), L.ptr(wU),
                                          stream), 'maint_rank1')
            L.check(lib.polar_maint_cache(p, N, L.ptr(ea), L.ptr(eb), L.ptr(coef), q, L.ptr(wX), L.ptr(wY),
                                          L.ptr(wU), L.ptr(h), L.ptr(t), stream), 'maint_cache')
            if want_trace:
                tr_hist.append(coef[:, 2 * q:2 * q + 2].clone())
        Fv = Fb
        L.check(lib.polar_rows_axpy(dcode, L.ptr(Fv), Fv.stride(0), q, a, b, L.ptr(gc_row), stream), 'rows_axpy')  # :250
        ph.mark('maintenance')
        op.add(a, b)                                                               # :251
        ph.mark('update')
        if progress is not None:
            progress('round', dict(round=rnd, edge=e, marginal=marg[-1], seconds=time.perf_counter() - t0,
                                   spmv=cnt.spmv))
    _sync()
    seconds = time.perf_counter() - t0
    if want_trace and not naive:
        for dn in tr_hist:
            dn_h = dn.cpu().numpy()
            traces.append(traces[-1] - math.fsum(om * dn_h[:, 1] / dn_h[:, 0]))
    gh = gaps_d[:len(seq)].cpu().numpy()
    with np.errstate(invalid='ignore', divide='ignore'):
        pgap = (gh[:, 0] - gh[:, 1]) / np.abs(gh[:, 0])
        rgap = (gh[:, 2] - gh[:, 3]) / np.abs(gh[:, 2])
    return dict(sequence=seq, seconds=seconds, spmv=cnt.spmv, rhs=cnt.rhs, p=p, rounds=len(seq),
                truncated=len(seq) < k, marginals=marg, trace_estimate=traces[:max(len(seq), 1)],
                solves=list(cnt.solves), precision='fp64' if dtype == torch.float64 else 'fp32',
                proxy_gap=pgap.tolist(), rerank_gap=rgap.tolist(), replace=replace,
                **({'phases': ph.acc} if profile else {}))


# ------------------------------------------------------------------------------------------------
def degree_greedy(G, cand, k):
    """polar_gpu.py:258-265 (host)."""
    deg = np.bincount(G.edges[:, 0], G.weights, G.n) + np.bincount(G.edges[:, 1], G.weights, G.n)
    live = np.ones(len(cand), bool)
    seq = []
    for _ in range(k):
        val = np.where(live, deg[cand[:, 0]] + deg[cand[:, 1]], np.inf)
        e = int(np.argmin(val))
        seq.append(e)
        live[e] = False
        deg[cand[e, 0]] += 1
        deg[cand[e, 1]] += 1
    return seq


def random_set(cand, k, seed):
    """polar_gpu.py:268-269 (host)."""
    return np.random.default_rng(seed).choice(len(cand), k, replace=False).tolist()


def true_marginals(G, cand, seq, s, obj='frac', eps=1e-9, tol=1e-10, kmax=None, poles=None,
                   precision='fp64', together=64):
    """Certified marginals along a sequence (polar_gpu.py:273-291): single-RHS multi-shift CG on
    the GPU with active=[a, b] and g = sum_j om_j ||u_j||^2 / (1 + b^T u_j).

    The reference solves position t after adding the edges seq[:t], one position at a time.  Here
    `together` positions at a time run as column groups of one operator: position t's column sees
    the added edges seq[:t] (edges before the chunk untagged, later ones as patch entries tagged
    with the groups of the positions after them, in addition order), so each column does the
    arithmetic of its own sequential solve, in one solve instead of `together`.  together=1 is the
    sequential loop."""
    L.require_gpu()
    dev = _device()
    kmax = kmax or len(seq)
    if poles is None:
        tau, om, qerr = globals()['poles'](obj, s, G.sigma, G.Lam(kmax), eps)
    else:
        tau, om = np.asarray(poles[0], float), np.asarray(poles[1], float)
        qerr = float('nan')
    track = obj != 'logdet'
    cnt = Counter()
    out = []
    ab = np.array([[int(cand[e, 0]), int(cand[e, 1])] for e in seq], dtype=np.int64).reshape(-1, 2)
    step = max(1, int(together))
    for c0 in range(0, len(ab), step):
        c1 = min(len(ab), c0 + step)
        m = c1 - c0
        op = Op(G, precision=precision)
        for i in range(c0):                                   # edges every position of the chunk sees
            op.add(ab[i, 0], ab[i, 1])
        if m == 1:
            a, b = int(ab[c0, 0]), int(ab[c0, 1])
            w = torch.zeros((G.n, 1), dtype=op.dtype, device=dev)
            w[a, 0] = 1.0
            w[b, 0] = -1.0
            U, hh, _ = msolve(op, w, tau, np.asarray([a, b]), tol, track=track, counter=cnt)[:3]
            bu = (U[:, 0, 0].double() - U[:, 1, 0].double()).cpu().numpy()[:, None]
            hh = hh.cpu().numpy() if track else None
        else:
            from . import batched
            for i in range(c0, c1 - 1):                       # edge i: the positions after it
                for t in range(i + 1, c1):
                    op.add(ab[i, 0], ab[i, 1], group=t - c0)
            act_h = np.unique(ab[c0:c1].ravel())
            slot = np.searchsorted(act_h, ab[c0:c1])
            cols = torch.arange(m, device=dev)
            W = torch.zeros((G.n, m), dtype=op.dtype, device=dev)
            W[torch.from_numpy(ab[c0:c1, 0]).to(dev), cols] = 1.0
            W[torch.from_numpy(ab[c0:c1, 1]).to(dev), cols] = -1.0
            act = torch.from_numpy(act_h.astype(np.int32)).to(dev)
            X, hb, it = batched.msolve_batched(op, batched.single_segment(G, len(act_h)), W, np.tile(tau, (m, 1)),
                                               act, tol, track=track, col_group=np.arange(m, dtype=np.int32))[:3]
            sa = torch.from_numpy(slot[:, 0]).to(dev)
            sb = torch.from_numpy(slot[:, 1]).to(dev)
            bu = (X[:, sa, cols].double() - X[:, sb, cols].double()).cpu().numpy()      # (p, m)
            hh = hb[:, 0, :].cpu().numpy() if track else None
            it_sum = int(np.asarray(it).sum())
            cnt.spmv += it_sum
            cnt.rhs += m
            cnt.solves.append(it_sum)
        for t in range(m):
            if obj == 'logdet':
                out.append(float(np.log1p(bu[0, t])))
            else:
                out.append(float(np.sum(om * hh[:, t] / (1 + bu[:, t]))))
    rho = tol * float(np.max((G.Lam(kmax) + tau) / (G.sigma + tau)))
    return np.array(out), dict(quadrature_bound=qerr, residual_rho=rho, p=len(tau), spmv=cnt.spmv)
