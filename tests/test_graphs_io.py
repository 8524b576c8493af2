"""Host-side graph I/O and generators (CPU)."""
import gzip
import json

import numpy as np

from polar_amd import configs, graphs


def test_snap_loader_preprocessing_and_cache(tmp_path, monkeypatch):
    """realworld/graphs.py:46-58 rule: relabel (all endpoints kept), undirect, dedup, no loops."""
    raw = tmp_path / 'raw'
    raw.mkdir()
    text = '# comment\n10 20\n20 10\n10 10\n30 20\n40 50\n'
    with gzip.open(raw / graphs.SNAP['ca-GrQc'], 'wt') as f:
        f.write(text)
    monkeypatch.setenv('POLAR_DATA', str(tmp_path))
    n, e, meta = graphs.load('ca-GrQc')
    assert n == 5 and e.tolist() == [[0, 1], [1, 2], [3, 4]]
    assert meta['m'] == 3 and len(meta['sha256']) == 64
    assert (tmp_path / 'npz' / 'ca-GrQc.npz').exists()
    n2, e2, meta2 = graphs.load('ca-GrQc')
    assert n2 == n and np.array_equal(e2, e) and meta2 == meta


def test_candidates_are_nonedges_and_reproducible():
    n = 300
    e = graphs.barabasi_albert(n, 3, 1)
    keys = set((e[:, 0] * n + e[:, 1]).tolist())
    for fn in (graphs.candidates, graphs.candidates_fast):
        c = fn(n, e, 500, 9)
        assert c.shape == (500, 2) and np.all(c[:, 0] < c[:, 1])
        assert not (set((c[:, 0] * n + c[:, 1]).tolist()) & keys)
        assert len(set(map(tuple, c.tolist()))) == 500
        assert np.array_equal(c, fn(n, e, 500, 9))
    c2 = graphs.candidates_2hop(n, e, 50, 3)
    assert c2.shape == (50, 2) and not (set((c2[:, 0] * n + c2[:, 1]).tolist()) & keys)
    assert len(graphs.all_nonedges(20, graphs.barabasi_albert(20, 2, 0))) > 0


def test_generators_match_their_specs():
    e = graphs.barabasi_albert(500, 5, 0)
    assert np.all(e[:, 0] < e[:, 1]) and len(np.unique(e, axis=0)) == len(e)
    cl = graphs.chung_lu(20000, 2.5, 20.0, 1)
    assert len(cl) == 200000 and np.all(cl[:, 0] < cl[:, 1]) and len(np.unique(cl, axis=0)) == len(cl)
    w = graphs.chung_lu_weights(20000, 2.5, 20.0)
    assert abs(w.mean() - 20.0) < 1e-9 and w[0] ** 2 <= w.sum() * (1 + 1e-9)
    rm = graphs.rmat(5000, 0.57, 0.19, 0.19, 0.05, 8, 100)
    assert len(rm) == 40000 and rm.max() < 5000 and np.all(rm[:, 0] < rm[:, 1])
    c = configs.make('c1')
    assert c['n'] == 2000 and c['N'] == 5000 and len(c['tau']) == 16 and c['q'] == 32
    assert c['m'] == len(c['edges']) and c['B'] == 1 and c['k'] == 10


def test_candidates_fast_is_uniform_over_nonedges():
    """candidates_fast draws a uniform N-subset of the non-edge pairs, like candidates
    (realworld/graphs.py:61-69): over 3000 seeds every non-edge is included with probability N/M
    (M non-edges; chi-square of the inclusion counts at the 0.1% level), no edge or loop is ever
    drawn, and the first row of the permuted pool is uniform over the non-edges too."""
    from scipy.stats import chi2
    n = 16
    e = graphs.barabasi_albert(n, 2, 3)
    ekeys = set((e[:, 0] * n + e[:, 1]).tolist())
    nonedges = [a * n + b for a in range(n) for b in range(a + 1, n) if a * n + b not in ekeys]
    M, N, trials = len(nonedges), 40, 3000
    col = {k: i for i, k in enumerate(nonedges)}
    for fn in (graphs.candidates, graphs.candidates_fast):
        inc = np.zeros(M)
        first = np.zeros(M)
        for seed in range(trials):
            c = fn(n, e, N, seed)
            k = c[:, 0] * n + c[:, 1]
            idx = [col[int(x)] for x in k]           # KeyError if an edge or a loop was drawn
            inc[idx] += 1
            first[idx[0]] += 1
        p = N / M
        stat = np.sum((inc - trials * p) ** 2 / (trials * p * (1 - p)))
        assert stat < chi2.ppf(0.999, M - 1), (fn.__name__, stat)
        exp = trials / M
        stat1 = np.sum((first - exp) ** 2 / exp)
        assert stat1 < chi2.ppf(0.999, M - 1), (fn.__name__, stat1)
