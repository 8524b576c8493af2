"""CPU FP64 oracle for the pole-shared Krylov core.  TEST INFRASTRUCTURE ONLY.

This module checks the product and is never part of it: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import it.  The product path (``polar_amd``) never imports it
and fails loudly when its HIP library is missing.

It restates, in NumPy/SciPy FP64, the semantics of the reference's ``realworld/polar_gpu.py``
(file:line citations below) with two additive extensions the HIP path also implements:

* ``msolve(..., freeze=True, trace=False)``: ``freeze=False`` reproduces the single-RHS variant
  of ``synthetic/scalability/benchmark.py:72-87`` (no pole freezing); ``trace=True`` adds a
  2-scalar per-pole tracker ``w^T (M+tau_j)^{-1} w`` (Gauss-quadrature identity of CG).
* ``polar(..., poles=(tau, om))`` pole override and per-round marginal / trace-estimate outputs.
* ``arith='cr'``: the same algorithm with a pinned arithmetic: every reduction (inner products,
  squared norms, row dots) is summed in double-double and rounded once (order-independent), the
  operator is applied as ((diag + shift) * x) - (sequential neighbour sum, rows longer than 64
  split in 64-entry chunks added in order) + added-edge terms in addition order, no fused
  multiply-add anywhere, and scalar sums over poles run in pole order.  The HIP path implements
  exactly this arithmetic, so its FP64 results equal the oracle's bit for bit (up to ties of the
  double-double rounding).  ``arith='numpy'`` (default) keeps NumPy's native reduction order,
  which reproduces the reference's own NumPy runs (fixture test).

Parity pins (what checks each function; tests/test_oracle.py):
* composite      -- closed form sum w/(x+t) vs x^-s within 256*4^-r, pole counts p=(K+2)r vs the
                    reference's stored time_*.json `p` fields; Gauss-Jacobi nodes vs SciPy.
* fixed_rule     -- closed form x^-s on the spectrum interval.
* Graph/Op       -- scipy dense (L + sigma I + sum b b^T) products.
* msolve         -- dense np.linalg.solve per pole; tracked norms vs ||u||^2; trace tracker vs
                    w^T (M+tau)^-1 w.
* polar          -- the reference's own fixture synthetic/scalability/results.json (sequences of
                    Naive and Inc-Full, spmv counts) and B=1 == Naive; dense closed-form greedy.
* true_marginals -- eigenvalue prefix gains.
* lanczos_quad   -- with m >= n steps the Gauss rule is exact: z^T f(A) z from the dense
                    eigendecomposition; slq_greedy -- its per-candidate estimates vs exact
                    z^T f(M + b b^T) z over the same probes (m = n).
"""
import math
import time

import numpy as np
import scipy.sparse as sp


# ------------------------------------------------------------------------------------------------
# Quadrature.  realworld/polar_gpu.py:18-36 (composite), :39-43 (poles).
def _jacobi_rule(beta, r):
    """r-point Gauss rule on [0,1] for the weight u^beta (polar_gpu.py:23-28): Golub-Welsch on the
    Jacobi matrix of the (0, beta) Jacobi weight, mapped by u=(1+x)/2."""
    i = np.arange(r, dtype=float)
    diag = np.empty(r)
    diag[0] = beta / (beta + 2)
    ii = i[1:]
    diag[1:] = beta * beta / ((2 * ii + beta) * (2 * ii + beta + 2))
    off = 2 / (2 * ii + beta) * np.sqrt(ii * ii * (ii + beta) ** 2
                                         / ((2 * ii + beta - 1) * (2 * ii + beta + 1)))
    x, Q = np.linalg.eigh(np.diag(diag) + np.diag(off, 1) + np.diag(off, -1))
    return (1 + x) / 2, Q[0] ** 2 / (beta + 1)


def composite(s, a, z, eps):
    """Positive composite Stieltjes rule for x^-s on [a, z] (polar_gpu.py:18-36).
    Returns (tau ascending, omega, bound 256*4^-r)."""
    r = int(np.ceil(np.log(256 / eps) / np.log(4)))
    K = int(np.ceil(np.log2(z / a)))
    cs = np.sin(np.pi * s) / np.pi
    lo, wlo = _jacobi_rule(-s, r)
    hi, whi = _jacobi_rule(s - 1, r)
    mid, wmid = np.polynomial.legendre.leggauss(r)
    ts = [a * lo]
    ws = [cs * a ** (1 - s) * wlo]
    for ell in range(K):
        L = a * 2 ** ell
        t = L * (3 + mid) / 2
        ts.append(t)
        ws.append(cs * L / 2 * wmid * t ** (-s))
    C = a * 2 ** K
    ts.append(C / hi)
    ws.append(cs * C ** (1 - s) * whi / hi)
    tau = np.concatenate(ts)
    om = np.concatenate(ws)
    o = np.argsort(tau)
    return tau[o], om[o], 256 * 4.0 ** (-r)


def fixed_rule(s, p, lo, hi):
    """Fixed-count p-pole Stieltjes rule for x^-s.  The Gauss-Jacobi rule of
    synthetic/scalability/benchmark.py:20-25 (t=(1-y)/(1+y)) rescaled by c=sqrt(lo*hi) so that its
    nodes straddle the spectrum [lo, hi]: tau=c*t, omega=c^(1-s)*w.  Returns (tau asc, omega, nan)."""
    i = np.arange(p)
    diag = (1 - 2 * s) / (4 * i * i - 1.0)
    j = np.arange(1, p)
    off = np.sqrt((j - s) * (j + s - 1)) / (2 * j - 1)
    off[0] *= np.sqrt(2)
    y, V = np.linalg.eigh(np.diag(diag) + np.diag(off, 1) + np.diag(off, -1))
    t = ((1 - y) / (1 + y))[::-1]
    w = (2 * V[0] ** 2 / (1 + y))[::-1]
    c = math.sqrt(lo * hi)
    return c * t, c ** (1 - s) * w, float('nan')


def poles(obj, s, sigma, Lam, eps):
    """polar_gpu.py:39-43: 'er'/'logdet' -> one pole at 0 with weight 1."""
    if obj in ('er', 'logdet'):
        return np.zeros(1), np.ones(1), 0.0
    return composite(s, sigma, Lam, eps)


# ------------------------------------------------------------------------------------------------
# Pinned-arithmetic helpers (arith='cr').
def ddsum(P, axis=0):
    """Double-double pairwise sum along `axis`, rounded once: the correctly rounded sum of the
    entries except in double-double rounding ties."""
    hi = np.moveaxis(np.array(P, dtype=np.float64), axis, 0)
    lo = np.zeros_like(hi)
    while hi.shape[0] > 1:
        if hi.shape[0] % 2:
            z = np.zeros((1,) + hi.shape[1:])
            hi = np.concatenate([hi, z])
            lo = np.concatenate([lo, z])
        ah, bh, al, bl = hi[0::2], hi[1::2], lo[0::2], lo[1::2]
        s_ = ah + bh
        bb = s_ - ah
        e = (ah - (s_ - bb)) + (bh - bb)
        e = e + (al + bl)
        hi = s_ + e
        lo = e - (hi - s_)
    return hi[0] + lo[0]


CHUNK = 64


class _Adjacency:
    """Off-diagonal adjacency in sorted CSR, for the pinned operator order."""

    def __init__(self, G):
        A = (sp.diags(G.M0.diagonal()) - G.M0).tocsr()
        A.eliminate_zeros()
        A.sort_indices()
        self.A = A
        self.diag = G.M0.diagonal().copy()
        deg = np.diff(A.indptr)
        self.long_rows = np.flatnonzero(deg > CHUNK)

    def neighbour_sum(self, x):
        acc = self.A @ x
        ip, ix, w = self.A.indptr, self.A.indices, self.A.data
        for r in self.long_rows:
            tot = None
            for c0 in range(ip[r], ip[r + 1], CHUNK):
                part = np.zeros(x.shape[1:])
                for k in range(c0, min(c0 + CHUNK, ip[r + 1])):
                    part = part + (w[k] * x[ix[k]] if w[k] != 1.0 else x[ix[k]])
                tot = part if tot is None else tot + part
            acc[r] = tot
        return acc


# ------------------------------------------------------------------------------------------------
class Graph:
    """M0 = diag(deg + sigma) - A as sorted CSR (polar_gpu.py:47-60)."""

    def __init__(self, n, edges, sigma, weights=None):
        self.n = int(n)
        self.sigma = float(sigma)
        edges = np.asarray(edges, dtype=np.int64)
        u, v = edges[:, 0], edges[:, 1]
        w = np.ones(len(u)) if weights is None else np.asarray(weights, float)
        deg = np.bincount(u, w, n) + np.bincount(v, w, n)
        A = sp.coo_matrix((np.r_[w, w], (np.r_[u, v], np.r_[v, u])), shape=(n, n)).tocsr()
        M = (sp.diags(deg + sigma) - A).tocsr()
        M.sort_indices()
        self.M0 = M
        self.dmax = float(deg.max())
        self.m = len(u)
        self.edges = edges
        self.weights = w

    def Lam(self, k):
        return self.sigma + 2 * self.dmax + 4 * k


class Op:
    """y = (M0 + shift I + sum_S b b^T) x with unit-weight added edges (polar_gpu.py:63-80)."""

    def __init__(self, G, arith='numpy'):
        self.M0 = G.M0
        self.A = []
        self.B = []
        self.arith = arith
        self._adj = _Adjacency(G) if arith == 'cr' else None

    def add(self, a, b):
        self.A.append(int(a))
        self.B.append(int(b))

    def __call__(self, x, shift):
        if self.arith == 'cr':
            xx = x if x.ndim == 2 else x[:, None]
            y = (self._adj.diag + shift)[:, None] * xx - self._adj.neighbour_sum(xx)
            for a, b in zip(self.A, self.B):
                y[a] = y[a] + (xx[a] - xx[b])
                y[b] = y[b] + (xx[b] - xx[a])
            return y if x.ndim == 2 else y[:, 0]
        y = self.M0 @ x
        if shift:
            y = y + shift * x
        if self.A:
            ia = np.asarray(self.A)
            ib = np.asarray(self.B)
            diff = x[ia] - x[ib]
            np.add.at(y, ia, diff)
            np.add.at(y, ib, -diff)
        return y


class Counter:
    """polar_gpu.py:83-86 (plus the per-solve list `solves` of column-iteration totals)."""

    def __init__(self):
        self.spmv = 0
        self.rhs = 0
        self.solves = []


def msolve(op, W, tau, active, tol, maxit=100000, track=False, counter=None, freeze=True,
           trace=False):
    colsum = ddsum if getattr(op, 'arith', 'numpy') == 'cr' else (lambda P: np.sum(P, axis=0))
    """Block multi-shift CG (Jegerlehner) for (M + tau_j I) U = W, all columns, all poles
    (polar_gpu.py:88-133).  Shifted iterates live only on the `active` rows.

    Returns (X[p, nA, R], h[p, R] or None, iters[R]) and, with trace=True, a 4th value
    e[p, R] = w_c^T (M + tau_j)^{-1} w_c from the recurrence e += a_j * g, g <- b_j * g
    (g = w^T p_j, using w^T r_t = 0 for t >= 1)."""
    W = np.asarray(W, dtype=float)
    if W.ndim == 1:
        W = W[:, None]
    n, R = W.shape
    tau = np.asarray(tau, dtype=float)
    p = len(tau)
    active = np.asarray(active)
    shift = (tau - tau[0])[:, None]                                   # :92
    r = W.copy()
    d = r.copy()
    rr = colsum(r * r)
    stop = tol * np.sqrt(rr)                                          # :93
    act = rr > 0                                                      # :94
    X = np.zeros((p, len(active), R))
    P = np.broadcast_to(W[active][None], (p, len(active), R)).copy()  # :95
    z = np.ones((p, R))
    zp = np.ones((p, R))
    live = np.broadcast_to(act[None], (p, R)).copy()                  # :96
    if track:
        h = np.zeros((p, R))
        ip = np.zeros((p, R))
        pn = np.broadcast_to(rr[None], (p, R)).copy()                 # :97-98
    if trace:
        te = np.zeros((p, R))
        tg = np.broadcast_to(rr[None], (p, R)).copy()
    ap = np.ones(R)
    bp = np.zeros(R)
    iters = np.zeros(R, dtype=np.int64)                               # :99
    with np.errstate(divide='ignore', invalid='ignore'):
        for it in range(maxit):                                       # :100
            v = op(d, float(tau[0]))
            dv = colsum(d * v)
            a = np.where(act, rr / np.where(act, dv, 1.0), 0.0)
            rn = r - a * v
            rrn = colsum(rn * rn)
            be = np.where(act, rrn / np.where(act, rr, 1.0), 0.0)
            c = np.where(act, a * bp / ap, 0.0)                       # :107
            zn = z / (1 + a * shift + c * (1 - z / zp))               # :108
            fac = zn / z
            m = live & act[None]                                      # :110
            aj = np.where(m, a * fac, 0.0)
            bj = be * fac * fac                                       # :111
            X += aj[:, None, :] * P                                   # :112
            P = np.where(m[:, None, :], zn[:, None, :] * rn[active][None] + bj[:, None, :] * P, P)
            if track:                                                 # :114-117
                h += 2 * aj * ip + aj * aj * pn
                ip = np.where(m, bj * (ip + aj * pn), ip)
                pn = np.where(m, zn * zn * rrn + bj * bj * pn, pn)
            if trace:
                te += aj * tg
                tg = np.where(m, bj * tg, tg)
            zp = np.where(m, z, zp)
            z = np.where(m, zn, z)                                    # :118
            nr = np.sqrt(rrn)
            if freeze:
                live &= ~(m & (np.abs(zn) * nr <= stop))              # :120
            iters += act
            done = act & (nr <= stop)                                 # :122
            d = np.where(act, rn + be * d, d)
            r = np.where(act, rn, r)
            rr = np.where(act, rrn, rr)
            ap = np.where(act, a, ap)
            bp = np.where(act, be, bp)                                # :123-124
            act &= ~done
            if it % 8 == 7 and not act.any():                         # :126
                break
        else:
            raise RuntimeError('multi-shift CG iteration cap')        # :128-129
    if counter is not None:
        counter.spmv += int(iters.sum())
        counter.rhs += R
        counter.solves.append(int(iters.sum()))
    out = (X, (h if track else None), iters)
    if trace:
        out = out + (te,)
    return out


# Column-parallel msolve for the CPU baseline: the columns of a block solve are independent
# (per-column CG scalars, per-column reductions), so splitting them over processes gives the same
# results as one call.  The operator's base matrix reaches the workers through fork (module state
# set before the pool starts); each task carries its added edges and its columns.
_POOL_OP = None


def _msolve_task(args):
    A, B, W, tau, active, tol, kw = args
    op = _POOL_OP
    op.A, op.B = list(A), list(B)
    return msolve(op, W, tau, active, tol, **kw)


def pmsolve(pool, nworkers, op, W, tau, active, tol, counter=None, **kw):
    """msolve over a process pool (same outputs as msolve): columns split into nworkers blocks."""
    R = W.shape[1]
    cuts = np.linspace(0, R, min(nworkers, R) + 1).astype(int)
    tasks = [(op.A, op.B, np.ascontiguousarray(W[:, a:b]), tau, active, tol, kw)
             for a, b in zip(cuts[:-1], cuts[1:])]
    parts = pool.map(_msolve_task, tasks)
    X = np.concatenate([o[0] for o in parts], axis=2)
    h = None if parts[0][1] is None else np.concatenate([o[1] for o in parts], axis=1)
    iters = np.concatenate([o[2] for o in parts])
    if counter is not None:
        counter.spmv += int(iters.sum())
        counter.rhs += R
        counter.solves.append(int(iters.sum()))
    out = (X, h, iters)
    if kw.get('trace'):
        out = out + (np.concatenate([o[3] for o in parts], axis=1),)
    return out


# ------------------------------------------------------------------------------------------------
# Sketches.  polar_gpu.py:137-165 ('block', 'gauss') plus the counter-hash Rademacher kind.
def _block_rows(rows, q, rng, center=False):
    """Stacked orthogonal q x q row blocks (polar_gpu.py:137-145)."""
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


def rademacher_signs(seed, stream, keys, q):
    """+-1 (float64) of shape (len(keys), q): sign = top bit of splitmix64(base + key*q + c) with
    base = splitmix64(seed*4 + stream).  Same definition as the HIP generator."""
    with np.errstate(over='ignore'):
        base = _splitmix64(np.uint64(seed * 4 + stream))
        k = np.asarray(keys, dtype=np.uint64)[:, None] * np.uint64(q) + np.arange(q, dtype=np.uint64)[None]
        hsh = _splitmix64((base + k) & _M64)
    return np.where((hsh >> np.uint64(63)) == 1, -1.0, 1.0)


def sketches(G, cand, q, seed, kind='block', chunk=4_000_000):
    """(Z n x q, F = C0^T U n x q, Gc N x q) as in polar_gpu.py:148-165.  kind='rademacher':
    entries +-1/sqrt(q) from the counter hash (streams 0: Z rows, 1: edge keys min*n+max,
    2: candidate rows); C0 has +1 at the smaller endpoint and -1 at the larger."""
    n = G.n
    if kind == 'rademacher':
        sq = 1.0 / math.sqrt(q)
        Z = rademacher_signs(seed, 0, np.arange(n), q) * sq
        F = np.zeros((n, q))
        u = np.minimum(G.edges[:, 0], G.edges[:, 1])
        v = np.maximum(G.edges[:, 0], G.edges[:, 1])
        sw = np.sqrt(G.weights)
        for lo in range(0, G.m, chunk):
            hi = min(G.m, lo + chunk)
            U = rademacher_signs(seed, 1, u[lo:hi] * n + v[lo:hi], q) * (sq * sw[lo:hi, None])
            np.add.at(F, u[lo:hi], U)
            np.add.at(F, v[lo:hi], -U)
        Gc = rademacher_signs(seed, 2, np.arange(len(cand)), q) * sq
        return Z, F, Gc
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
    """polar_gpu.py:169-173."""
    per_col = 8.0 * (10 * n + 4 * p * nA)
    cap = 16 if n > 3e6 else 64
    return int(max(8, min(cap, budget // per_col)))


# ------------------------------------------------------------------------------------------------
def polar(G, cand, k, s, q=128, B=16, eps=1e-2, tol=1e-7, obj='frac', sketch=None, naive=False,
          seed=2026, kind='block', trace=None, time_limit=None, poles=None, width=None,
          freeze=True, arith='numpy', workers=1):
    """POLAR greedy with cache backend (polar_gpu.py:180-254).

    Extensions: `poles=(tau, om)` overrides the rule; `width` overrides chunk_width; the result
    also carries `marginals` (per-round G_q(e*;S) = sum_j om_j ||u_j||^2/(1+b^T u_j) of the winner;
    the winner's proxy score for naive), `trace_estimate` (sketched sum_j om_j tr(Z^T(M_S+tau_j)^-1 Z)
    after each round, index 0 = before any addition) and `bank_seconds`.  workers > 1 splits the
    columns of every block solve over that many forked processes (same results; used to time the
    CPU baseline on a many-core host)."""
    if workers > 1:
        import multiprocessing
        global _POOL_OP
        _POOL_OP = Op(G, arith=arith)
        pool = multiprocessing.get_context('fork').Pool(workers)
        try:
            return _polar(G, cand, k, s, q, B, eps, tol, obj, sketch, naive, seed, kind, trace, time_limit,
                          poles, width, freeze, arith, (pool, workers))
        finally:
            pool.terminate()
            _POOL_OP = None
    return _polar(G, cand, k, s, q, B, eps, tol, obj, sketch, naive, seed, kind, trace, time_limit,
                  poles, width, freeze, arith, None)


def _polar(G, cand, k, s, q, B, eps, tol, obj, sketch, naive, seed, kind, trace, time_limit, poles, width,
           freeze, arith, par):
    def solve(*a, **kw):
        return pmsolve(par[0], par[1], *a, **kw) if par else msolve(*a, **kw)
    n, N = G.n, len(cand)
    sigma = G.sigma
    if poles is None:
        tau, om, _ = globals()['poles'](obj, s, sigma, G.Lam(k), eps)
    else:
        tau, om = np.asarray(poles[0], float), np.asarray(poles[1], float)
    p = len(tau)
    if sketch is None:
        sketch = sketches(G, cand, q, seed, kind)
    Z, F, Gc = sketch
    q = Z.shape[1]
    F = np.array(F, dtype=float, copy=True)
    cand = np.asarray(cand)
    active = np.unique(cand)                                          # :191
    es = np.searchsorted(active, cand)
    ea, eb = es[:, 0], es[:, 1]
    tau_c = tau[:, None]
    om_c = om[:, None]
    op = Op(G, arith=arith)
    cr = arith == 'cr'
    cnt = Counter()
    live = np.ones(N, bool)
    seq, marg, traces = [], [], []

    def qsum(P):                      # reduction over the sketch axis (last)
        return ddsum(P, axis=-1) if cr else np.sum(P, axis=-1)

    def polesum(T):                   # sum over poles (axis 0), in pole order
        if not cr:
            return np.sum(T, axis=0)
        acc = np.zeros(T.shape[1:])
        for j in range(T.shape[0]):
            acc = acc + T[j]
        return acc
    t0 = time.perf_counter()
    if width is None:
        width = chunk_width(n, p, len(active))                        # :198
    want_trace = obj != 'logdet'

    def bank():                                                       # :200-206
        RHS = np.concatenate([Z, F], axis=1)
        parts, terms = [], []
        for lo in range(0, 2 * q, width):
            out = solve(op, RHS[:, lo:lo + width], tau, active, tol, counter=cnt,
                        freeze=freeze, trace=want_trace)
            parts.append(out[0])
            if want_trace:
                ncol = min(max(q - lo, 0), out[0].shape[2])
                terms.append((om_c * out[3][:, :ncol]).ravel())
        XY = np.concatenate(parts, axis=2)
        X, Y = XY[:, :, :q].copy(), XY[:, :, q:].copy()
        tr = math.fsum(np.concatenate(terms)) if terms else 0.0
        h = qsum((X[:, ea] - X[:, eb]) ** 2)
        t = qsum((Y[:, ea] - Y[:, eb]) ** 2)
        return X, Y, h, t, tr

    def proxy(h, t):                                                  # :208-211
        if obj == 'logdet':
            return np.log1p(t[0] + sigma * h[0])
        return polesum(om_c * h / (1 + t + (sigma + tau_c) * h))

    X, Y, h, t, tr = bank()
    bank_seconds = time.perf_counter() - t0
    traces.append(tr if want_trace else None)
    for rnd in range(k):
        if rnd and time_limit is not None and time.perf_counter() - t0 > time_limit:  # :215
            break
        if naive and rnd:
            X, Y, h, t, tr = bank()
            traces[-1] = tr if want_trace else None
        sc = proxy(h, t)
        sc[~live] = -np.inf                                           # :219
        Br = 1 if naive else min(B, int(live.sum()))
        short = np.argsort(-sc, kind='stable')[:Br]                   # :221
        if naive:
            e = int(short[0])
            marg.append(float(sc[e]))
        else:
            Wb = np.zeros((n, Br))
            Wb[cand[short, 0], np.arange(Br)] = 1.0
            Wb[cand[short, 1], np.arange(Br)] = -1.0                  # :225-226
            need_track = obj != 'logdet'
            U, hh, _ = solve(op, Wb, tau, active, tol, track=need_track, counter=cnt,
                             freeze=freeze)[:3]
            ar = np.arange(Br)
            bu = U[:, es[short, 0], ar] - U[:, es[short, 1], ar]      # :230
            if obj == 'logdet':
                gsc = np.log1p(bu[0])
            else:
                gsc = polesum(om_c * hh / (1 + bu))                   # :236
            j = 0 if Br == 1 else int(np.argmax(gsc))
            e = int(short[j])
            u = U[:, :, j]
            marg.append(float(gsc[j]))
        if trace is not None:
            trace.append(dict(shortlist=[int(x) for x in short], chosen=e, proxy=sc[short].tolist(),
                              scores=[] if naive else np.asarray(gsc, float).tolist()))
        seq.append(e)
        live[e] = False
        aa, bb = es[e]
        a, b = cand[e]
        if not naive and rnd + 1 < k:                                  # :242-249
            xi = X[:, aa] - X[:, bb]
            eta = Gc[e][None] - (Y[:, aa] - Y[:, bb])
            D = 1 + u[:, aa] - u[:, bb]
            cc = (u[:, ea] - u[:, eb]) / D[:, None]
            if cr:
                wX = ddsum(X * xi[:, None, :], axis=-1)
                wY = ddsum(Y * eta[:, None, :], axis=-1)
            else:
                wX = np.einsum('pnq,pq->pn', X, xi)
                wY = np.einsum('pnq,pq->pn', Y, eta)
            nxi = qsum(xi * xi)
            neta = qsum(eta * eta)
            h = np.maximum(0, h - 2 * cc * (wX[:, ea] - wX[:, eb]) + cc * cc * nxi[:, None])
            t = np.maximum(0, t + 2 * cc * (wY[:, ea] - wY[:, eb]) + cc * cc * neta[:, None])
            X -= u[:, :, None] * xi[:, None, :] / D[:, None, None]
            Y += u[:, :, None] * eta[:, None, :] / D[:, None, None]
            if want_trace:
                traces.append(traces[-1] - math.fsum(om * nxi / D))
        elif naive and want_trace:
            traces.append(None)
        F[a] += Gc[e]                                                 # :250
        F[b] -= Gc[e]
        op.add(a, b)                                                  # :251
    if naive and want_trace:
        traces = traces[:len(seq)]
    return dict(sequence=seq, seconds=time.perf_counter() - t0, spmv=cnt.spmv, rhs=cnt.rhs, p=p,
                rounds=len(seq), truncated=len(seq) < k, marginals=marg,
                trace_estimate=traces, solves=list(cnt.solves), bank_seconds=bank_seconds)


# ------------------------------------------------------------------------------------------------
def degree_greedy(G, cand, k):
    """polar_gpu.py:258-265."""
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
    """polar_gpu.py:268-269."""
    return np.random.default_rng(seed).choice(len(cand), k, replace=False).tolist()


def true_marginals(G, cand, seq, s, obj='frac', eps=1e-9, tol=1e-10, kmax=None, poles=None,
                   arith='numpy'):
    """Certified marginals along a sequence (polar_gpu.py:273-291)."""
    kmax = kmax or len(seq)
    if poles is None:
        tau, om, qerr = globals()['poles'](obj, s, G.sigma, G.Lam(kmax), eps)
    else:
        tau, om = np.asarray(poles[0], float), np.asarray(poles[1], float)
        qerr = float('nan')
    op = Op(G, arith=arith)
    cnt = Counter()
    out = []
    for e in seq:
        a, b = int(cand[e, 0]), int(cand[e, 1])
        w = np.zeros((G.n, 1))
        w[a, 0] = 1.0
        w[b, 0] = -1.0
        U, hh, _ = msolve(op, w, tau, np.asarray([a, b]), tol, track=obj != 'logdet', counter=cnt)
        bu = U[:, 0, 0] - U[:, 1, 0]
        if obj == 'logdet':
            g = float(np.log1p(bu[0]))
        else:
            g = float(np.sum(om * hh[:, 0] / (1 + bu)))
        out.append(g)
        op.add(a, b)
    rho = tol * float(np.max((G.Lam(kmax) + tau) / (G.sigma + tau)))
    return np.array(out), dict(quadrature_bound=qerr, residual_rho=rho, p=len(tau), spmv=cnt.spmv)


# ------------------------------------------------------------------------------------------------
# Dense evaluators (realworld/exact.py:12-17, 65-82), used to pin the oracle.
def frac_value(lam, s, obj):
    if obj == 'frac':
        return float(np.sum(lam ** (-s)))
    if obj == 'er':
        return float(np.sum(1 / lam))
    return -float(np.sum(np.log(lam)))


def eval_prefix(M0, cand, seq, s, obj='frac'):
    M = np.array(M0, dtype=float, copy=True)
    base = frac_value(np.linalg.eigvalsh(M), s, obj)
    out = []
    for e in seq:
        a, b = int(cand[e, 0]), int(cand[e, 1])
        M[a, a] += 1
        M[b, b] += 1
        M[a, b] -= 1
        M[b, a] -= 1
        out.append(base - frac_value(np.linalg.eigvalsh(M), s, obj))
    return np.array(out)


# ------------------------------------------------------------------------------------------------
# SLQ-Greedy baseline (realworld/slq.py:10-58): Hutchinson probes shared by the base matrix and
# every candidate (common random numbers), m-step Lanczos quadrature for z^T f(A) z.
def lanczos_quad(apply, Z, m, f):
    """z_c^T f(A_c) z_c for every column c of Z (n x C) by m-step Lanczos (slq.py:10-21): plain
    three-term recurrence, no reorthogonalisation; T_c's eigenpairs give the Gauss rule."""
    n, C = Z.shape
    nz = np.sqrt(np.sum(Z * Z, 0))
    V = Z / nz
    Vp = np.zeros_like(V)
    al = np.zeros((m, C))
    be = np.zeros((m, C))
    b = np.zeros(C)
    for j in range(m):
        W = apply(V) - b * Vp
        a = np.sum(W * V, 0)
        W = W - a * V
        b = np.sqrt(np.sum(W * W, 0))
        al[j] = a
        be[j] = b
        Vp = V
        V = W / np.where(b > 0, b, 1.0)
    T = np.zeros((C, m, m))
    idx = np.arange(m)
    T[:, idx, idx] = al.T
    T[:, idx[:-1], idx[1:]] = be[:-1].T
    T[:, idx[1:], idx[:-1]] = be[:-1].T
    th, U = np.linalg.eigh(T)
    return nz ** 2 * np.sum(U[:, 0, :] ** 2 * f(np.maximum(th, 1e-300)), axis=1)


def slq_greedy(G, cand, k, s, probes=20, m=40, chunk=4000, seed=7):
    """slq.py:24-58: per round, base = mean_z z^T f(M_S) z and for every live candidate the mean
    of z^T f(M_S + b b^T) z over the same probes; the argmax of base - that is selected."""
    f = lambda x: x ** (-s)
    n = G.n
    rng = np.random.default_rng(seed)
    live = np.ones(len(cand), bool)
    seq = []
    op = Op(G)
    spmv = 0
    ests = []
    for _ in range(k):
        Z = rng.choice([-1.0, 1.0], size=(n, probes))
        base = float(np.mean(lanczos_quad(lambda X: op(X, 0.0), Z, m, f)))
        spmv += probes * m
        ids = np.flatnonzero(live)
        est = np.empty(len(ids))
        per = max(1, chunk // probes)
        for lo in range(0, len(ids), per):
            sub = ids[lo:lo + per]
            c = len(sub)
            ea = np.repeat(cand[sub, 0], probes)
            eb = np.repeat(cand[sub, 1], probes)
            cols = np.arange(c * probes)

            def apply(X):
                Y = op(X, 0.0)
                d = X[ea, cols] - X[eb, cols]
                Y[ea, cols] += d
                Y[eb, cols] -= d
                return Y
            q = lanczos_quad(apply, np.tile(Z, (1, c)), m, f).reshape(c, probes)
            est[lo:lo + c] = base - np.mean(q, 1)
            spmv += c * probes * m
        e = int(ids[np.argmax(est)])
        seq.append(e)
        ests.append(est)
        live[e] = False
        op.add(int(cand[e, 0]), int(cand[e, 1]))
    return dict(sequence=seq, spmv=spmv, probes=probes, lanczos_steps=m, estimates=ests)
