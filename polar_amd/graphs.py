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


# ------------------------------------------------------------------------------------------------
# Dataset loading (realworld/graphs.py:9-58): SNAP / KONECT edge lists under $POLAR_DATA/raw,
# relabelled, undirected, deduplicated, self-loops removed, cached as npz.  The SNAP/KONECT
# files are not shipped; synthetic workloads use the generators above (see configs.make).
import gzip as _gzip
import hashlib as _hashlib
import io as _io
import json as _json
import os as _os
import tarfile as _tarfile
from pathlib import Path as _Path

SNAP = {
    'email-Eu-core': 'email-Eu-core.txt.gz', 'ca-GrQc': 'ca-GrQc.txt.gz', 'ca-HepPh': 'ca-HepPh.txt.gz',
    'ca-CondMat': 'ca-CondMat.txt.gz', 'email-Enron': 'email-Enron.txt.gz',
    'com-DBLP': 'com-dblp.ungraph.txt.gz', 'com-Amazon': 'com-amazon.ungraph.txt.gz',
    'web-Google': 'web-Google.txt.gz', 'roadNet-PA': 'roadNet-PA.txt.gz',
    'com-Youtube': 'com-youtube.ungraph.txt.gz', 'as-Skitter': 'as-skitter.txt.gz',
    'roadNet-CA': 'roadNet-CA.txt.gz', 'com-LiveJournal': 'com-lj.ungraph.txt.gz',
}
KONECT = {
    'Zachary': 'download.tsv.ucidata-zachary.tar.bz2', 'Dolphins': 'download.tsv.dolphins.tar.bz2',
    'LesMis': 'download.tsv.moreno_lesmis.tar.bz2', 'Polbooks': 'download.tsv.dimacs10-polbooks.tar.bz2',
    'Hamster': 'download.tsv.petster-hamster-household.tar.bz2',
}


def data_dir():
    return _Path(_os.environ.get('POLAR_DATA', _Path.cwd() / 'data'))


def _parse(text_bytes, comment):
    rows = []
    for line in _io.BytesIO(text_bytes).read().decode(errors='replace').splitlines():
        line = line.strip()
        if not line or line.startswith(comment):
            continue
        parts = line.split()
        if len(parts) >= 2:
            rows.append((int(parts[0]), int(parts[1])))
    return np.asarray(rows, dtype=np.int64).reshape(-1, 2)


def preprocess(raw):
    """Relabel to 0..n-1 (all endpoint vertices kept), undirect, dedup, drop self-loops."""
    ids, inv = np.unique(np.asarray(raw).ravel(), return_inverse=True)
    e = inv.reshape(-1, 2)
    e = np.sort(e, axis=1)
    e = e[e[:, 0] != e[:, 1]]
    e = np.unique(e, axis=0)
    return len(ids), e.astype(np.int64)


def load(name):
    """(n, edges, meta) from $POLAR_DATA/raw with an npz cache in $POLAR_DATA/npz."""
    root = data_dir()
    npz = root / 'npz'
    cache = npz / f'{name}.npz'
    if cache.exists():
        z = np.load(cache, allow_pickle=False)
        return int(z['n']), z['edges'], _json.loads(str(z['meta']))
    raw_dir = root / 'raw'
    if name in SNAP:
        path = raw_dir / SNAP[name]
        raw = _parse(_gzip.open(path).read(), '#')
    elif name in KONECT:
        path = raw_dir / KONECT[name]
        with _tarfile.open(path, 'r:bz2') as tf:
            mem = next(m for m in tf.getmembers() if _Path(m.name).name.startswith('out.'))
            raw = _parse(tf.extractfile(mem).read(), '%')
    else:
        raise FileNotFoundError(f'unknown graph {name!r} (not a SNAP/KONECT name or synthetic config)')
    n, e = preprocess(raw)
    meta = dict(name=name, source=path.name, sha256=_hashlib.sha256(path.read_bytes()).hexdigest(),
                n=int(n), m=int(len(e)),
                preprocessing='retain all endpoint vertices; undirect; deduplicate; remove self-loops')
    npz.mkdir(parents=True, exist_ok=True)
    np.savez(cache, n=n, edges=e, meta=_json.dumps(meta))
    return n, e, meta


def all_nonedges(n, edges):
    """Every non-edge pair u<v (realworld/graphs.py:72-74), for tiny graphs."""
    present = set(map(tuple, np.asarray(edges).tolist()))
    return np.array([(u, v) for u in range(n) for v in range(u + 1, n) if (u, v) not in present],
                    dtype=np.int64).reshape(-1, 2)


def candidates_2hop(n, edges, N, seed):
    """Friend-of-friend pool by random 2-step walks (realworld/graphs.py:77-93)."""
    import scipy.sparse as sp
    A = sp.coo_matrix((np.ones(2 * len(edges)), (np.r_[edges[:, 0], edges[:, 1]], np.r_[edges[:, 1], edges[:, 0]])),
                      shape=(n, n)).tocsr()
    ptr, nbr = A.indptr, A.indices
    deg = np.diff(ptr)
    keys = np.sort(edges[:, 0] * n + edges[:, 1])
    rng = np.random.default_rng(seed)
    C = set()
    starts = np.flatnonzero(deg > 0)
    while len(C) < N:
        u = int(starts[rng.integers(len(starts))])
        v = int(nbr[ptr[u] + rng.integers(deg[u])])
        w = int(nbr[ptr[v] + rng.integers(deg[v])])
        a, b = min(u, w), max(u, w)
        i = np.searchsorted(keys, a * n + b)
        if a != b and not (i < len(keys) and keys[i] == a * n + b):
            C.add((a, b))
    c = np.asarray(sorted(C), dtype=np.int64)
    return np.ascontiguousarray(c[rng.permutation(N)])
