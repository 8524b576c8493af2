"""Operators of a million rows and more: the single-system operator kernel then runs with a larger
block budget (a launch shape only), so its results must equal, bit for bit, those of the
1,024-block shape that the smaller parity tests cover; the product is also checked against SciPy."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from polar_amd import graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402

N_ROWS = 1_100_000


@pytest.fixture(scope='module')
def big_graph():
    e = graphs.chung_lu(N_ROWS, 2.5, 4.0, 11)
    return pg.Graph(N_ROWS, e, 0.5)


def test_large_operator_product_matches_scipy(big_graph):
    G = big_graph
    rng = np.random.default_rng(0)
    X = rng.standard_normal((G.n, 8))
    Y = pg.Op(G)(torch.from_numpy(X).cuda(), 0.3).cpu().numpy()
    want = G.M0.get() @ X + 0.3 * X
    assert np.max(np.abs(Y - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_large_operator_launch_shape_is_bitwise_neutral(big_graph, precision):
    G = big_graph
    rng = np.random.default_rng(1)
    dt = torch.float64 if precision == 'fp64' else torch.float32
    W = torch.from_numpy(rng.standard_normal((G.n, 8))).to(dt).cuda()
    tau = np.array([0.0, 0.1, 1.0, 10.0])
    act = np.arange(0, G.n, 7)
    out = []
    for grid_max in (0, 1024):
        op = pg.Op(G, precision=precision)
        op.grid_max = grid_max
        X, h, it = pg.msolve(op, W, tau, act, 1e-8, track=True)
        out.append((X.cpu(), h.cpu(), it))
    (X0, h0, i0), (X1, h1, i1) = out
    assert np.array_equal(i0, i1) and i0.max() > 10
    assert torch.equal(X0, X1) and torch.equal(h0, h1)
