"""Dense `effect` evaluator (polar_amd/exact.py; realworld/exact.py:65-82) on the GPU (rocSOLVER
eigvalsh) against the oracle's NumPy restatement and the twin closed form: an edge between two
isolated vertices turns their eigenvalues (sigma, sigma) into (sigma, sigma + 2), so its prefix gain
is sigma^-s - (sigma+2)^-s (frac), 1/sigma - 1/(sigma+2) (er) and log(sigma+2) - log(sigma)
(logdet) whatever the rest of the graph (realworld/exact.py:20-28)."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import exact, graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _graph_with_isolated(n=700, iso=6, seed=3):
    e = graphs.barabasi_albert(n - iso, 3, seed)           # vertices n-iso .. n-1 stay isolated
    return n, e


def test_eval_prefix_matches_oracle_and_twin_closed_form():
    n, e = _graph_with_isolated()
    sigma = 0.1
    G = pg.Graph(n, e, sigma)
    M0 = G.M0.get().toarray()
    cand = np.concatenate([graphs.candidates(n - 6, e, 40, 5), [[n - 6, n - 5], [n - 4, n - 3]]])
    seq = [3, 40, 7, 41, 12]                               # two twin pairs among ordinary edges
    for obj, s in (('frac', 0.5), ('frac', 0.25), ('er', 1.0), ('logdet', 0.0)):
        got = exact.eval_prefix(M0, cand, seq, s, obj)
        ref = O.eval_prefix(M0, cand, seq, s, obj)
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-10, obj
        step = np.diff(np.r_[0.0, got])
        twin = {'frac': sigma ** -s - (sigma + 2) ** -s, 'er': 1 / sigma - 1 / (sigma + 2),
                'logdet': np.log(sigma + 2) - np.log(sigma)}[obj]
        for pos in (1, 3):                                 # the rounds that added a twin pair
            assert abs(step[pos] - twin) <= 1e-9 * abs(twin), (obj, pos, step[pos], twin)


def test_eval_prefix_multi_equals_single_targets():
    n, e = _graph_with_isolated(500, 4, 8)
    G = pg.Graph(n, e, 0.3)
    M0 = G.M0.get().toarray()
    cand = graphs.candidates(n - 4, e, 30, 2)
    seq = [0, 5, 9, 2]
    targets = [('frac', 0.25), ('frac', 0.75), ('er', 1.0), ('logdet', 0.0)]
    multi = exact.eval_prefix_multi(M0, cand, seq, targets)
    for (obj, s), v in zip(targets, multi):
        ref = O.eval_prefix(M0, cand, seq, s, obj)
        assert np.max(np.abs(v - ref) / np.abs(ref)) <= 1e-10, (obj, s)
