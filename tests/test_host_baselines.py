"""Host-side baselines and helpers of the drop-in (polar_gpu.py:83-86, 169-173, 258-269), pinned by
hand-worked cases: degree-greedy picks the live candidate with the smallest endpoint-degree sum
(ties: lowest index) and adds 1 to both endpoints; random_set draws k distinct candidates from
the seed; chunk_width is the reference's memory rule."""
import numpy as np
import pytest

pytest.importorskip('torch')

from polar_amd import polar_gpu as pg  # noqa: E402


def test_degree_greedy_hand_worked():
    # path 0-1-2-3: degrees 1, 2, 2, 1
    G = pg.Graph(4, np.array([[0, 1], [1, 2], [2, 3]]), 0.1)
    cand = np.array([[0, 2], [1, 3], [0, 3]])
    # (0,3) has sum 2; then (0,2) and (1,3) both have sum 4: the lower index wins
    assert pg.degree_greedy(G, cand, 2) == [2, 0]
    assert pg.degree_greedy(G, cand, 3) == [2, 0, 1]


def test_degree_greedy_uses_weighted_degree():
    G = pg.Graph(4, np.array([[0, 1], [2, 3]]), 0.1, weights=np.array([5.0, 0.5]))
    cand = np.array([[0, 2], [1, 3], [2, 3]])
    # weighted degrees 5, 5, 0.5, 0.5: (2,3) has sum 1, then (0,2) and (1,3) both 5 + 1.5
    assert pg.degree_greedy(G, cand, 3) == [2, 0, 1]


def test_random_set_is_distinct_seeded_and_in_range():
    a = pg.random_set(np.zeros((50, 2), int), 10, 7)
    assert len(set(a)) == 10 and all(0 <= x < 50 for x in a)
    assert a == pg.random_set(np.zeros((50, 2), int), 10, 7)
    assert a != pg.random_set(np.zeros((50, 2), int), 10, 8)
    assert a == np.random.default_rng(7).choice(50, 10, replace=False).tolist()


@pytest.mark.parametrize('n,p,nA,want', [(4_000_000, 32, 3_973_097, 8),    # floor of 8 columns
                                         (1000, 16, 500, 64),              # cap 64 below 3e6 rows
                                         (3_500_000, 1, 10, 16),           # cap 16 above 3e6 rows
                                         (1_000_000, 5, 1_000_000, 41)])   # 1e10 // (8 (1e7 + 2e7))
def test_chunk_width_rule(n, p, nA, want):
    assert pg.chunk_width(n, p, nA) == want


def test_counter_fields():
    c = pg.Counter()
    assert (c.spmv, c.rhs) == (0, 0)
