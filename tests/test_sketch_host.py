"""Host sketches (polar_gpu.py:137-165) of the drop-in against the oracle's, and the properties
that fix them: block rows are stacked orthonormal q x q blocks (Z centred), F = C0^T U sums to
zero over the vertices (every edge adds U to one endpoint and subtracts it from the other), and
Rademacher entries are +-1/sqrt(q)."""
import numpy as np
import pytest

pytest.importorskip('torch')

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _graph(weighted):
    n = 260
    e = graphs.barabasi_albert(n, 3, 4)
    w = np.random.default_rng(1).uniform(0.5, 1.5, len(e)) if weighted else None
    cand = graphs.candidates(n, e, 90, 2)
    return n, e, w, cand


@pytest.mark.parametrize('kind', ['block', 'gauss', 'rademacher'])
@pytest.mark.parametrize('weighted', [False, True])
def test_host_sketches_equal_oracle(kind, weighted):
    n, e, w, cand = _graph(weighted)
    a = O.sketches(O.Graph(n, e, 0.2, weights=w), cand, 12, 7, kind=kind)
    b = pg.sketches(pg.Graph(n, e, 0.2, weights=w), cand, 12, 7, kind=kind)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize('kind', ['block', 'gauss', 'rademacher'])
def test_sketch_properties(kind):
    n, e, w, cand = _graph(True)
    q = 12
    Z, F, Gc = pg.sketches(pg.Graph(n, e, 0.2, weights=w), cand, q, 3, kind=kind)
    Z, F, Gc = (np.asarray(x) for x in (Z, F, Gc))
    assert Z.shape == (n, q) and F.shape == (n, q) and Gc.shape == (len(cand), q)
    assert np.max(np.abs(F.sum(0))) < 1e-12 * max(1.0, np.max(np.abs(F)))
    if kind == 'block':
        assert np.max(np.abs(Z.mean(0))) < 1e-14
        for k in range(len(cand) // q):
            blk = Gc[k * q:(k + 1) * q]
            assert np.allclose(blk @ blk.T, np.eye(q), atol=1e-12)
    if kind == 'rademacher':
        assert np.allclose(np.abs(Z), 1 / np.sqrt(q)) and np.allclose(np.abs(Gc), 1 / np.sqrt(q))
        # F row of a vertex = sum over its edges of +-sqrt(w) * (+-1/sqrt(q))
        deg_w = np.zeros(n)
        np.add.at(deg_w, e[:, 0], np.sqrt(w))
        np.add.at(deg_w, e[:, 1], np.sqrt(w))
        assert np.all(np.abs(F) <= deg_w[:, None] / np.sqrt(q) + 1e-12)


def test_device_rademacher_sketch_rejects_multigraphs():
    """The device Rademacher sketch hashes each merged adjacency entry once; the host form adds one
    row per input edge.  They agree on simple graphs only, so a repeated edge or a self-loop is
    refused on the device path (before any GPU call) and the host path still accepts it."""
    from polar_amd import polar_gpu as pg
    e = np.array([[0, 1], [1, 2], [2, 3]])
    cand = np.array([[0, 2], [1, 3]])
    assert pg.Graph(4, e, 0.1).simple
    for bad in (np.r_[e, [[1, 0]]], np.r_[e, [[2, 2]]]):
        G = pg.Graph(4, bad, 0.1)
        assert not G.simple
        for mode in (True, 'lazy'):
            with pytest.raises(ValueError, match='simple graph'):
                pg.sketches(G, cand, 8, 3, kind='rademacher', on_device=mode)
        Z, F, Gc = pg.sketches(G, cand, 8, 3, kind='rademacher')
        assert F.shape == (4, 8)
