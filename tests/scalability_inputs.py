"""Inputs of the reference's CPU scalability fixture (synthetic/scalability/results.json).

The fixture stores selected-edge sequences and spmv counts for seeded synthetic graphs.  To check
the oracle (and the HIP path) against it, the seeded inputs are regenerated here with the same
NumPy RNG draw order as synthetic/scalability/benchmark.py:26-71 (ring-random and preferential
topologies, 'local'/'global' candidate supports, Gaussian sketches Z, F=C^T U, G).  The draw order
is a data format that has to match for the fixture to apply; the code is written independently.
"""
import numpy as np

P_POLES, Q_SKETCH, K_ROUNDS, N_CAND, TOL, SIGMA, S_EXP = 20, 24, 15, 3000, 1e-7, 1.0, 0.5


def quad(s, r):
    """Unscaled Gauss-Jacobi Stieltjes rule of benchmark.py:20-25 (nodes ascending)."""
    i = np.arange(r)
    diag = (1 - 2 * s) / (4 * i * i - 1.0)
    j = np.arange(1, r)
    off = np.sqrt((j - s) * (j + s - 1)) / (2 * j - 1)
    off[0] *= np.sqrt(2)
    y, V = np.linalg.eigh(np.diag(diag) + np.diag(off, 1) + np.diag(off, -1))
    return ((1 - y) / (1 + y))[::-1].copy(), (2 * V[0] ** 2 / (1 + y))[::-1].copy()


def _ring_random(n, rng):
    ring = np.stack([np.arange(n), np.roll(np.arange(n), -1)], axis=1)
    extra = rng.integers(n, size=(2 * n, 2))
    e = np.sort(np.concatenate([ring, extra]), axis=1)
    return np.unique(e[e[:, 0] != e[:, 1]], axis=0)


def _small_world(n, rng):
    ring = {(min(i, (i + h) % n), max(i, (i + h) % n)) for i in range(n) for h in (1, 2, 3)}
    for i in range(n):
        for h in (1, 2, 3):
            old = (min(i, (i + h) % n), max(i, (i + h) % n))
            if rng.random() < .1:
                ring.remove(old)
                while True:
                    j = int(rng.integers(n))
                    new = (min(i, j), max(i, j))
                    if j != i and new not in ring:
                        break
                ring.add(new)
    return np.array(sorted(ring), dtype=np.int64)


def _preferential(n, rng):
    ed = [(i, j) for i in range(4) for j in range(i + 1, 4)]
    urn = [x for e in ed for x in e]
    for v in range(4, n):
        picked = set()
        while len(picked) < 3:
            picked.add(urn[int(rng.integers(len(urn)))])
        for u in sorted(picked):
            ed.append((u, v))
            urn.extend((u, v))
    return np.array(ed, dtype=np.int64)


def make(n, kind, support):
    """Returns dict(edges, cand, Z, F, G) for one fixture configuration (seed 7+n)."""
    rng = np.random.default_rng(7 + n)
    if kind == 'ring-random':
        edges = _ring_random(n, rng)
    elif kind == 'small-world':
        edges = _small_world(n, rng)
    elif kind == 'preferential':
        edges = _preferential(n, rng)
    else:
        raise ValueError(kind)
    a, b = edges[:, 0], edges[:, 1]
    blocked = set(map(tuple, edges.tolist()))
    if support == 'local':
        pool = np.array([(i, j) for i in range(96) for j in range(i + 1, 96) if (i, j) not in blocked])
        cand = pool[rng.choice(len(pool), N_CAND, replace=False)]
    else:
        chosen = set()
        while len(chosen) < N_CAND:
            i, j = sorted(map(int, rng.integers(n, size=2)))
            if i != j and (i, j) not in blocked:
                chosen.add((i, j))
        cand = np.array(sorted(chosen), dtype=np.int64)
    Z = rng.normal(size=(n, Q_SKETCH)) / np.sqrt(Q_SKETCH)
    U = rng.normal(size=(len(a), Q_SKETCH)) / np.sqrt(Q_SKETCH)
    F = np.zeros((n, Q_SKETCH))
    np.add.at(F, a, U)
    np.add.at(F, b, -U)
    G = rng.normal(size=(N_CAND, Q_SKETCH)) / np.sqrt(Q_SKETCH)
    return dict(n=n, edges=edges.astype(np.int64), cand=np.asarray(cand, dtype=np.int64), Z=Z, F=F, G=G)
