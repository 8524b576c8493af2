"""Synthetic graph generators and candidate pools (host, NumPy).

Every generated edge list follows the reference's preprocessing rule (realworld/graphs.py:52-53):
undirected, each edge stored as (min, max), duplicates and self-loops removed, int64.

* barabasi_albert(n, m0, seed)          -- preferential attachment, m0 edges per arriving vertex
* chung_lu(n, gamma, mean_deg, seed, m)  -- expected-degree power law; endpoints drawn ~ w until m
                                            unique edges; w_i = (i + i0)^(-1/(gamma-1)) with i0
                                            chosen so that max w^2 = sum w (Chung-Lu admissibility)
* rmat(n, a, b, c, d, edge_factor, seed) -- recursive-matrix edges on 2^ceil(log2 n), ids >= n dropped
* candidates(n, edges, N, seed)          -- realworld/graphs.py:61-69 draw order (exact)
* candidates_fast(n, edges, N, seed)     -- same distribution, vectorised (for N up to 1e7)
"""
import numpy as np


def simple_edges(e, n=None):
    e = np.asarray(e, dtype=np.int64).reshape(-1, 2)
    e = np.sort(e, axis=1)
    e = e[e[:, 0] != e[:, 1]]
    if len(e) == 0:
        return e
    if n is None:
        n = int(e.max()) + 1
    key = np.unique(e[:, 0] * n + e[:, 1])
    return np.stack([key // n, key % n], axis=1)


def barabasi_albert(n, m0, seed):
    rng = np.random.default_rng(seed)
    edges = []
    repeated = []
    targets = list(range(m0))
    for v in range(m0, n):
        for t in targets:
            edges.append((t, v))
        repeated.extend(targets)
        repeated.extend([v] * m0)
        chosen = set()
        while len(chosen) < m0:
            chosen.add(repeated[int(rng.integers(len(repeated)))])
        targets = sorted(chosen)
    return simple_edges(np.array(edges, dtype=np.int64), n)


def chung_lu_weights(n, gamma, mean_deg):
    ex = 1.0 / (gamma - 1.0)
    i0 = 1.0
    for _ in range(60):
        w = (np.arange(n) + i0) ** (-ex)
        w *= mean_deg * n / w.sum()
        if w[0] ** 2 <= w.sum() * (1 + 1e-9):
            break
        i0 *= 1.5
    lo, hi = i0 / 1.5 if i0 > 1 else 0.0, i0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        w = (np.arange(n) + mid) ** (-ex)
        w *= mean_deg * n / w.sum()
        if w[0] ** 2 <= w.sum():
            hi = mid
        else:
            lo = mid
    w = (np.arange(n) + hi) ** (-ex)
    return w * (mean_deg * n / w.sum())


def chung_lu(n, gamma, mean_deg, seed, m=None):
    """Sample endpoint pairs i.i.d. proportional to w until m unique undirected non-loop edges
    (default m = round(n*mean_deg/2)); a surplus of distinct pairs is thinned uniformly."""
    rng = np.random.default_rng(seed)
    w = chung_lu_weights(n, gamma, mean_deg)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    m = int(round(n * mean_deg / 2)) if m is None else int(m)
    keys = np.empty(0, dtype=np.int64)
    factor = 1.15
    while len(keys) < m:
        b = int((m - len(keys)) * factor) + 1024
        u = np.searchsorted(cdf, rng.random(b), side='right').astype(np.int64)
        v = np.searchsorted(cdf, rng.random(b), side='right').astype(np.int64)
        np.minimum(u, n - 1, out=u)
        np.minimum(v, n - 1, out=v)
        lo = np.minimum(u, v)
        hi = np.maximum(u, v)
        ok = lo != hi
        keys = np.unique(np.concatenate([keys, lo[ok] * n + hi[ok]]))
        factor *= 1.5
    if len(keys) > m:
        keys = np.sort(rng.choice(keys, m, replace=False))
    return np.stack([keys // n, keys % n], axis=1)


def rmat(n, a, b, c, d, edge_factor, seed):
    rng = np.random.default_rng(seed)
    scale = int(np.ceil(np.log2(max(n, 2))))
    m = int(edge_factor * n)
    probs = np.array([a, b, c, d], dtype=float)
    probs /= probs.sum()
    cum = np.cumsum(probs)
    out = []
    total = 0
    while total < m:
        cnt = int((m - total) * 1.25) + 1024
        u = np.zeros(cnt, dtype=np.int64)
        v = np.zeros(cnt, dtype=np.int64)
        for lvl in range(scale):
            qd = np.searchsorted(cum, rng.random(cnt), side='right')
            u |= ((qd >> 1) & 1).astype(np.int64) << lvl
            v |= (qd & 1).astype(np.int64) << lvl
        ok = (u < n) & (v < n)
        e = simple_edges(np.stack([u[ok], v[ok]], axis=1), n)
        out.append(e)
        allv = simple_edges(np.concatenate(out), n)
        out = [allv]
        total = len(allv)
        if cnt > 50 * m:
            break
    e = out[0]
    if len(e) > m:
        e = e[np.sort(rng.choice(len(e), m, replace=False))]
    return e


def candidates(n, edges, N, seed):
    """Uniform non-edge pool, the reference's exact draw order (realworld/graphs.py:61-69)."""
    keys = np.sort(edges[:, 0] * n + edges[:, 1])
    rng = np.random.default_rng(seed)
    C = set()
    while len(C) < N:
        u, v = sorted(map(int, rng.integers(n, size=2)))
        i = np.searchsorted(keys, u * n + v)
        if u != v and not (i < len(keys) and keys[i] == u * n + v):
            C.add((u, v))
    c = np.asarray(sorted(C), dtype=np.int64)
    return np.ascontiguousarray(c[rng.permutation(N)])


def candidates_fast(n, edges, N, seed, batch=1 << 23):
    """Same distribution as `candidates` (uniform over non-edge unordered pairs), vectorised:
    batches of uniform pairs, drop loops / edges / repeats, first N distinct in draw order, then
    sorted and permuted as the reference does."""
    ekeys = np.sort(edges[:, 0].astype(np.int64) * n + edges[:, 1])
    rng = np.random.default_rng(seed)
    got = np.empty(0, dtype=np.int64)
    while len(got) < N:
        b = min(batch, 2 * (N - len(got)) + 1024)
        uv = rng.integers(n, size=(b, 2))
        lo = uv.min(1).astype(np.int64)
        hi = uv.max(1).astype(np.int64)
        k = lo * n + hi
        k = k[lo != hi]
        pos = np.searchsorted(ekeys, k)
        pos = np.minimum(pos, max(len(ekeys) - 1, 0))
        if len(ekeys):
            k = k[ekeys[pos] != k]
        allk = np.concatenate([got, k])
        _, first = np.unique(allk, return_index=True)
        got = allk[np.sort(first)]
    got = np.sort(got[:N])
    c = np.stack([got // n, got % n], axis=1)
    return np.ascontiguousarray(c[rng.permutation(N)])
