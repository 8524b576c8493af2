"""Row runs of the operator kernel: a work unit of several consecutive rows walked as one entry
sequence (a gather batch spans row boundaries).  Each row still sums its entries in adjacency
order, so every result must equal the one-row-per-unit walk bit for bit:
  * products on a graph with long rows (skipped inside runs), isolated rows (empty rows inside and
    at the ends of runs), added edges (row patch), unit and general weights, FP64 and FP32, widths
    from 3 to 128 columns and run lengths up to the row group's lane count;
  * multi-shift solves (the fused d.v partials; FP32 with the blocked update, residual
    replacement and column compaction), the batched form (runs inside segments) and column groups;
  * the FP64 greedy against the oracle with every operator using runs;
  * the default: runs of 16 rows on graphs of 2^20 rows and more with mean degree <= 8."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import batched, configs, graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def gapped_hub_graph(weighted=False, seed=5):
    """Power-law graph with two hubs of more than 256 neighbours, relabelled v -> v + v // 50 so
    that every 51st row is isolated (empty rows at many positions inside runs)."""
    n0 = 3000
    e = graphs.chung_lu(n0, 2.3, 8.0, seed)
    hubs = np.array([[0, j] for j in range(1, 900)] + [[7, j] for j in range(300, 1000)])
    e = graphs.simple_edges(np.concatenate([e, hubs]), n0)
    e = e + e // 50
    n = n0 + n0 // 50 + 3                      # three isolated rows at the end too
    w = np.random.default_rng(seed).uniform(0.5, 1.5, len(e)) if weighted else None
    return pg.Graph(n, e, 0.3, weights=w)


def _product(G, precision, R, run, X, pairs):
    op = pg.Op(G, precision=precision)
    op.run_rows = run
    for a, b in pairs:
        op.add(a, b)
    return op(X, 0.0125)


@pytest.mark.parametrize('weighted', [False, True])
@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('R', [3, 8, 24, 128])
def test_spmm_row_runs_bitwise(weighted, precision, R):
    G = gapped_hub_graph(weighted)
    assert len(G.schedule()['long_rows']) >= 2 and np.sum(np.diff(G._indptr) == 0) >= 60
    dt = torch.float64 if precision == 'fp64' else torch.float32
    X = torch.from_numpy(np.random.default_rng(R).standard_normal((G.n, R))).to('cuda', dt)
    pairs = [(3, 2500), (0, 3002), (7, 8), (50, 51)]
    ref = _product(G, precision, R, 1, X, pairs)
    for run in (2, 5, 16, 63):
        got = _product(G, precision, R, run, X, pairs)
        assert torch.equal(got, ref), run
    M = G.M0.get().toarray()
    for a, b in pairs:
        M[a, a] += 1
        M[b, b] += 1
        M[a, b] -= 1
        M[b, a] -= 1
    Xh = X.double().cpu().numpy()
    want = M @ Xh + 0.0125 * Xh
    scale = np.abs(M).sum(1, keepdims=True) * np.abs(Xh).max()
    tol = 1e-12 if precision == 'fp64' else 2e-6
    assert np.max(np.abs(ref.double().cpu().numpy() - want) / scale) < tol


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_msolve_row_runs_bitwise(precision):
    """Solves with the fused d.v partials; FP32 with the blocked update and residual replacement, and
    zero columns so that column compaction happens."""
    G = gapped_hub_graph()
    dt = torch.float64 if precision == 'fp64' else torch.float32
    rng = np.random.default_rng(3)
    W = rng.standard_normal((G.n, 16))
    W[:, ::3] = 0.0
    W = torch.from_numpy(W).to('cuda', dt)
    tau = np.array([0.0, 0.05, 0.5, 5.0])
    act = np.arange(0, G.n, 3)
    kw = dict(track=True, trace=True)
    if precision == 'fp32':
        kw.update(block=True, replace=True)
    out = []
    for run in (1, 7):
        op = pg.Op(G, precision=precision)
        op.run_rows = run
        X, h, it, te = pg.msolve(op, W, tau, act, 1e-9, **kw)
        out.append((X.cpu(), h.cpu(), it, te.cpu()))
    (X0, h0, i0, t0), (X1, h1, i1, t1) = out
    assert np.array_equal(i0, i1) and i0.max() > 24
    assert torch.equal(X0, X1) and torch.equal(h0, h1) and torch.equal(t0, t1)


def test_batched_and_column_groups_with_row_runs(monkeypatch):
    """Runs inside segments of a block-diagonal union, and column groups: every operator with runs
    of 5 rows equals the one-row walk."""
    e3 = graphs.chung_lu(1500, 2.5, 6.0, 4)
    Gs = [pg.Graph(400, graphs.barabasi_albert(400, 3, 7), 0.1),
          gapped_hub_graph(),
          pg.Graph(1500, e3, 0.2, weights=np.random.default_rng(1).uniform(0.5, 1.5, len(e3)))]
    p, R = 6, 8
    poles = [pg.fixed_rule(s, p, G.sigma, G.Lam(4))[:2] for s, G in zip([0.25, 0.5, 0.75], Gs)]
    tau = np.stack([t for t, _ in poles])
    rng = np.random.default_rng(4)
    W = torch.from_numpy(rng.standard_normal((sum(G.n for G in Gs), R))).cuda()
    res = []
    for run in (0, 5):
        monkeypatch.setattr(pg.Graph, 'run_rows', lambda self, r=run: r)
        U = batched.UnionGraph(Gs)
        a0 = np.r_[0, np.cumsum([G.n for G in Gs])]
        segs = batched.Segments(*U.segments(a0))
        active = torch.arange(U.n, dtype=torch.int32).cuda()
        res.append(batched.msolve_batched(pg.Op(U), segs, W, tau, active, 1e-9, track=True))
    (X0, h0, i0), (X1, h1, i1) = [r[:3] for r in res]
    assert np.array_equal(i0, i1) and torch.equal(X0, X1) and torch.equal(h0, h1)
    # column groups: three systems on one graph (FP32-mixed, shared-seed bank, replacement)
    G = gapped_hub_graph()
    cand = graphs.candidates(G.n, G.edges, 400, 12)
    s_list = [0.25, 0.5, 0.75]
    poles = [pg.fixed_rule(s, 8, G.sigma, G.Lam(5))[:2] for s in s_list]
    outs = []
    for run in (0, 3):
        monkeypatch.setattr(pg.Graph, 'run_rows', lambda self, r=run: r)
        sk = pg.sketches(G, cand, 16, 21, kind='rademacher', on_device='lazy')
        outs.append(batched.polar_columns(G, cand, 4, s_list, q=16, B=3, sketch=sk, poles=poles, precision='fp32'))
    for a, b in zip(*outs):
        for key in ('sequence', 'marginals', 'trace_estimate', 'spmv', 'rhs'):
            assert a[key] == b[key], key


def test_polar_fp64_with_row_runs_equals_oracle(monkeypatch):
    """c1 (BA n=2000) with every operator of the greedy using runs of 7 rows: the oracle's
    sequence, counters, marginals and trace estimates."""
    monkeypatch.setattr(pg.Graph, 'run_rows', lambda self: 7)
    c = configs.make('c1')
    Go = O.Graph(c['n'], c['edges'], c['sigma'])
    sk = O.sketches(Go, c['cand'], c['q'], c['seed'], kind='rademacher')
    kw = dict(q=c['q'], B=4, tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']))
    ro = O.polar(Go, c['cand'], 6, c['s'], width=64, arith='cr', **kw)
    rg = pg.polar(pg.Graph(c['n'], c['edges'], c['sigma']), c['cand'], 6, c['s'], **kw)
    assert rg['sequence'] == ro['sequence'] and rg['spmv'] == ro['spmv']
    rel = np.max(np.abs(np.array(rg['marginals']) - ro['marginals']) / np.abs(ro['marginals']))
    assert rel < 1e-13
    rel = np.max(np.abs(np.array(rg['trace_estimate']) - ro['trace_estimate']) / np.abs(ro['trace_estimate']))
    assert rel < 1e-13


def test_default_row_runs():
    lat = 1100
    idx = np.arange(lat * lat).reshape(lat, lat)
    e = np.concatenate([np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1),
                        np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], 1)])
    assert pg.Graph(lat * lat, e, 0.1).run_rows() == 16          # 1.21M rows, mean degree < 4
    assert pg.Graph(2000, graphs.barabasi_albert(2000, 2, 0), 0.1).run_rows() == 0   # small
