"""Edge cases of the HIP path against the oracle (FP64 pinned arithmetic: exact counters, 1e-13
relative values): odd and single-column widths, one pole, zero columns, isolated vertices,
exhausted candidate pools, B larger than the pool, widths that need several block solves."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _isolated_graph(n=400, seed=3):
    """A BA core on the first half of the vertices; the second half is isolated."""
    e = graphs.barabasi_albert(n // 2, 2, seed)
    return n, e


@pytest.mark.parametrize('R', [1, 3, 5, 7, 130])
@pytest.mark.parametrize('p', [1, 5])
def test_msolve_odd_widths_and_single_pole(R, p):
    n, e = _isolated_graph()
    Go = O.Graph(n, e, 0.3)
    G = pg.Graph(n, e, 0.3)
    rng = np.random.default_rng(R * 7 + p)
    W = rng.standard_normal((n, R))
    W[:, 0] = 0.0                                           # a column that is never active
    tau = np.sort(rng.uniform(0.0, 2.0, p))
    act = np.sort(rng.choice(n, n // 3, replace=False))
    Xo, ho, ito = O.msolve(O.Op(Go, arith='cr'), W, tau, act, 1e-9, track=True)[:3]
    Xg, hg, itg = pg.msolve(pg.Op(G), torch.from_numpy(W).cuda(), tau, act, 1e-9, track=True)
    assert np.array_equal(ito, itg) and ito[0] == 0
    assert _rel(Xg.cpu().numpy(), Xo) <= 1e-13 and _rel(hg.cpu().numpy(), ho) <= 1e-13


def test_polar_isolated_vertices_exhausted_pool_and_large_B():
    """Pool of 12 candidates, k = 12 rounds (the pool runs out), B = 40 > N: the shortlist is every
    live candidate; pairs of isolated vertices have the closed-form gain sigma^-s - (sigma+2)^-s."""
    n, e = _isolated_graph(60, 5)
    sigma, s = 0.25, 0.5
    cand = graphs.candidates(n, e, 12, 8)
    Go = O.Graph(n, e, sigma)
    G = pg.Graph(n, e, sigma)
    tau, om, _ = pg.fixed_rule(s, 24, sigma, G.Lam(12))
    sk = O.sketches(Go, cand, 16, 4, kind='rademacher')
    kw = dict(q=16, B=40, tol=1e-10, sketch=sk, poles=(tau, om))
    ro = O.polar(Go, cand, 12, s, arith='cr', **kw)
    rg = pg.polar(G, cand, 12, s, **kw)
    assert rg['sequence'] == ro['sequence'] and sorted(rg['sequence']) == list(range(12))
    assert (rg['spmv'], rg['rhs']) == (ro['spmv'], ro['rhs'])
    assert _rel(rg['marginals'], ro['marginals']) <= 1e-12
    # a pair of isolated vertices: u_j = (e_a - e_b)/(sigma + tau_j), so its re-rank marginal is
    # f(sigma) - f(sigma + 2) with f(x) = sum_j om_j/(x + tau_j), the rule's x^-s (while no added
    # edge touches a or b)
    rule = lambda x: float(np.sum(om / (x + tau)))
    touched = set()
    checked = 0
    for r, e_ in enumerate(rg['sequence']):
        a, b = (int(v) for v in cand[e_])
        if min(a, b) >= n // 2 and a not in touched and b not in touched:
            assert abs(rg['marginals'][r] - (rule(sigma) - rule(sigma + 2))) <= 1e-12 * rg['marginals'][r]
            checked += 1
        touched |= {a, b}
    assert checked >= 1


def test_polar_wide_bank_split_into_block_solves():
    """2q = 160 > 128 FP64 columns per block solve: the bank runs as several solves; same result
    as the oracle (columns are independent)."""
    n = 500
    e = graphs.barabasi_albert(n, 3, 2)
    cand = graphs.candidates(n, e, 60, 3)
    Go, G = O.Graph(n, e, 0.2), pg.Graph(n, e, 0.2)
    tau, om, _ = pg.fixed_rule(0.5, 6, 0.2, G.Lam(3))
    sk = O.sketches(Go, cand, 80, 6, kind='rademacher')
    kw = dict(q=80, B=2, tol=1e-9, sketch=sk, poles=(tau, om))
    ro = O.polar(Go, cand, 3, 0.5, arith='cr', **kw)
    rg = pg.polar(G, cand, 3, 0.5, width=48, **kw)
    assert rg['sequence'] == ro['sequence'] and (rg['spmv'], rg['rhs']) == (ro['spmv'], ro['rhs'])
    assert _rel(rg['marginals'], ro['marginals']) <= 1e-12
    assert _rel(rg['trace_estimate'], ro['trace_estimate']) <= 1e-12


@pytest.mark.parametrize('q,p', [(5, 3), (7, 40), (7, 70)])
def test_polar_odd_sketch_width_parity(q, p):
    """Odd q (the rank-1 pass, the pair gathers and the bank run their scalar tail columns); the
    rank-1 pass runs row tiles of 4 passes (p = 3), of one pass (p = 40) and, at p = 70, whose
    tile does not fit its LDS budget, the pole-major walk: same sequence, counters, marginals and
    traces as the oracle."""
    n = 300
    e = graphs.barabasi_albert(n, 3, 7)
    sigma, s = 0.2, 0.5
    cand = graphs.candidates(n, e, 120, 5)
    Go, G = O.Graph(n, e, sigma), pg.Graph(n, e, sigma)
    tau, om, _ = pg.fixed_rule(s, p, sigma, G.Lam(5))
    sk = O.sketches(Go, cand, q, 9, kind='rademacher')
    kw = dict(q=q, B=3, tol=1e-9, sketch=sk, poles=(tau, om))
    ro = O.polar(Go, cand, 5, s, arith='cr', **kw)
    rg = pg.polar(G, cand, 5, s, **kw)
    assert rg['sequence'] == ro['sequence'] and (rg['spmv'], rg['rhs']) == (ro['spmv'], ro['rhs'])
    assert _rel(rg['marginals'], ro['marginals']) <= 1e-12
    assert _rel(rg['trace_estimate'], ro['trace_estimate']) <= 1e-12
