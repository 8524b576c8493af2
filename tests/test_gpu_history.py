"""History re-rank solves (POLAR_HISTORY): the solve keeps every residual and every pole's per-step
coefficients, and the shifted solutions X_j are formed afterwards only where the re-rank needs
them: on the two endpoint rows of every shortlisted column and on every active row of the
winner's column.

* The CG itself is the blocked FP32 solve's: iteration counts, tracked norms and trace terms are
  equal bit for bit (the finisher only records more).
* X_j formed from the history equals the window-flushed X_j within the FP32 rounding of the
  window flush (relative to the column's largest entry, 1e-5), with and without residual
  replacement, with column groups; in FP64 the two agree to 1e-12.
* polar(), polar_columns(): the history re-rank selects the same edges as the flushed re-rank,
  with marginals and trace estimates within 1e-6 (FP32; 1e-11 in FP64 with the blocked update),
  and a history too short for the solve falls back to the flushed solve with the same results.
* The mode is used with the blocked update only (never on the pinned FP64 path) and where at least
  256 iterations fit in the history budget (a share of the available device memory); once a
  history solve of a greedy run outruns its history, the run's later re-rank solves are flushed.
"""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from polar_amd import batched, configs, graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _graph(n=3000, seed=5):
    e = graphs.chung_lu(n, 2.3, 8.0, seed)
    hubs = np.array([[0, j] for j in range(1, 400)])
    return pg.Graph(n, graphs.simple_edges(np.concatenate([e, hubs]), n), 0.2)


def _rhs(G, R, seed):
    """Re-rank-like right-hand sides e_a - e_b."""
    rng = np.random.default_rng(seed)
    W = np.zeros((G.n, R))
    for c in range(R):
        a, b = rng.choice(G.n, 2, replace=False)
        W[a, c], W[b, c] = 1.0, -1.0
    return W


@pytest.mark.parametrize('precision,replace', [('fp32', False), ('fp32', True), ('fp64', False)])
def test_history_solve_matches_flushed_solve(precision, replace):
    G = _graph()
    R, tol = 12, 1e-9
    dt = torch.float32 if precision == 'fp32' else torch.float64
    W = torch.from_numpy(_rhs(G, R, 1)).to('cuda', dt)
    tau = np.array([0.0, 0.03, 0.3, 3.0, 30.0, 300.0])
    act_h = np.arange(0, G.n, 2)
    act = torch.from_numpy(act_h.astype(np.int32)).cuda()
    op = pg.Op(G, precision=precision)
    X, h, it, te = pg.msolve(op, W, tau, act, tol, track=True, trace=True, block=True, replace=replace)
    _, h2, it2, te2, H = pg.msolve(op, W, tau, act, tol, track=True, trace=True, block=True, replace=replace,
                                   history=512)
    assert np.array_equal(it, it2) and it.max() > 40
    assert torch.equal(h, h2) and torch.equal(te, te2)
    bound = 1e-5 if precision == 'fp32' else 1e-12
    Xd = X.double()
    scale = Xd.abs().amax(dim=1, keepdim=True)                    # per (pole, column)
    # every column on every active row
    Xh = H.cols(torch.arange(R, dtype=torch.int32, device='cuda'), act).double()   # (R, p, nA)
    err = (Xh.permute(1, 2, 0) - Xd).abs() / scale
    assert float(err.max()) < bound
    # three chosen rows of every column (absolute row ids among the active rows)
    rng = np.random.default_rng(2)
    slots = rng.integers(0, len(act_h), size=(3, R))
    rows = torch.from_numpy(act_h[slots].astype(np.int32)).cuda()
    E = H.rows(rows).double()                                     # (p, 3, R)
    want = torch.stack([Xd[:, torch.from_numpy(slots[k]).cuda(), torch.arange(R, device='cuda')] for k in range(3)], 1)
    assert float(((E - want).abs() / scale.squeeze(1)[:, None, :]).max()) < bound


def test_history_with_column_groups_matches_flushed_solve():
    G = _graph()
    R, p = 12, 6
    cg = np.repeat(np.arange(3), 4).astype(np.int32)
    tau = np.stack([np.geomspace(1e-3, 100.0, p) * f for f in (1.0, 1.5, 2.0)])
    segs = batched.single_segment(G, G.n)
    act = torch.arange(G.n, dtype=torch.int32, device='cuda')
    W = torch.from_numpy(_rhs(G, R, 3)).float().cuda()
    op = pg.Op(G, precision='fp32')
    X, h, it = batched.msolve_batched(op, segs, W, tau, act, 1e-8, track=True, block=True, replace=True,
                                      col_group=cg)[:3]
    _, h2, it2, H = batched.msolve_batched(op, segs, W, tau, act, 1e-8, track=True, block=True, replace=True,
                                           col_group=cg, history=512)
    assert np.array_equal(it, it2) and torch.equal(h, h2)
    Xh = H.cols(torch.arange(R, dtype=torch.int32, device='cuda'), act).double().permute(1, 2, 0)
    Xd = X.double()
    assert float(((Xh - Xd).abs() / Xd.abs().amax(dim=1, keepdim=True)).max()) < 1e-5


def test_history_cap_reached_raises_iteration_cap():
    G = _graph()
    W = torch.from_numpy(_rhs(G, 4, 4)).float().cuda()
    act = torch.arange(G.n, dtype=torch.int32, device='cuda')
    with pytest.raises(RuntimeError, match='iteration cap'):
        pg.msolve(pg.Op(G, precision='fp32'), W, np.array([0.0, 1.0]), act, 1e-12, block=True, history=16)


def _greedy_pair(fn, monkeypatch):
    """fn() with history re-rank solves, with flushed ones, and with a history cap too short."""
    out = [fn()]
    monkeypatch.setattr(pg, 'history_cap', lambda n, R, dtype, block=None: 0)
    out.append(fn())
    monkeypatch.setattr(pg, 'history_cap', lambda n, R, dtype, block=None: 8)
    out.append(fn())
    return out


def _same(a, b, bound=1e-6):
    assert a['sequence'] == b['sequence']
    assert (a['spmv'], a['rhs']) == (b['spmv'], b['rhs'])
    m_a, m_b = np.array(a['marginals']), np.array(b['marginals'])
    assert np.max(np.abs(m_a - m_b) / np.abs(m_b)) < bound
    if a['trace_estimate'][0] is None:          # logdet forms no trace estimate
        assert b['trace_estimate'][0] is None
        return
    t_a, t_b = np.array(a['trace_estimate']), np.array(b['trace_estimate'])
    assert np.max(np.abs(t_a - t_b) / np.abs(t_b)) < bound


def test_history_cap_rule():
    f32, f64 = torch.float32, torch.float64
    big = 250e9
    assert pg.history_cap(100_000, 24, f32) == pg.HISTORY_STEPS       # the device's own free memory
    # c3's re-rank (n=1M, 16 columns) needs about 420 iterations: on an idle device (90% of its
    # memory available, whatever earlier tests hold now) the budget keeps more than twice that
    idle = 0.9 * torch.cuda.mem_get_info()[1]
    assert pg.history_cap(1_000_000, 16, f32, avail=idle) > 840
    assert pg.history_cap(100_000, 24, f32, avail=big) == pg.HISTORY_STEPS
    assert pg.history_cap(100_000, 24, f64, avail=big) == 0           # the pinned FP64 path
    assert pg.history_cap(100_000, 24, f64, block=True, avail=big) == pg.HISTORY_STEPS
    assert pg.history_cap(100_000, 24, f32, block=False, avail=big) == 0
    # c3's re-rank (n=1M, 16 columns): the budget, not the step limit, sets the cap
    budget = min(pg.HISTORY_BYTES, int(pg.HISTORY_SHARE * big))
    assert pg.history_cap(1_000_000, 16, f32, avail=big) == budget // (1_000_000 * 16 * 4) > 900
    # c5's re-rank (n=4M, 16 columns) beside its caches: fewer than 256 steps fit
    assert pg.history_cap(4_000_000, 16, f32, avail=60e9) == 0
    assert pg.history_cap(1_000_000, 16, f32, avail=8e9) == 0


def test_polar_history_rerank_equals_flushed_rerank(monkeypatch):
    c = configs.make('c2', 0.02)
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    sk = pg.sketches(G, c['cand'], 16, c['seed'], kind='rademacher', on_device='lazy')
    assert pg.history_cap(G.n, 4, torch.float32) > 0

    def run():
        return pg.polar(G, c['cand'], 6, c['s'], q=16, B=4, tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']),
                        precision='fp32')
    hist, flushed, capped = _greedy_pair(run, monkeypatch)
    _same(hist, flushed)
    _same(capped, flushed)


@pytest.mark.parametrize('precision,bound', [('fp32', 1e-6), ('fp64', 1e-11)])
def test_polar_columns_history_rerank_equals_flushed_rerank(monkeypatch, precision, bound):
    """FP32, and FP64 with the blocked update (the bench's FP64 leg on the FP32 schedule)."""
    c = configs.make('c2', 0.02)
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    s_list = [0.25, 0.5, 0.75]
    poles = [pg.fixed_rule(s, 8, G.sigma, G.Lam(5))[:2] for s in s_list]
    dt = torch.float32 if precision == 'fp32' else torch.float64
    assert pg.history_cap(G.n, 9, dt, block=True) > 0

    def run():
        sk = pg.sketches(G, c['cand'], 16, c['seed'], kind='rademacher', on_device='lazy')
        return batched.polar_columns(G, c['cand'], 5, s_list, q=16, B=3, sketch=sk, poles=poles, precision=precision,
                                     block=True)
    hist, flushed, capped = _greedy_pair(run, monkeypatch)
    for a, b, d in zip(hist, flushed, capped):
        _same(a, b, bound)
        _same(d, b, bound)


@pytest.mark.parametrize('freeze', [True, False])
def test_history_with_zero_columns_and_graph_flag(freeze):
    """Zero right-hand sides (converged before the first iteration), the hipGraph flag (ignored by
    history solves) and freezing on or off: the same counts and norms as the flushed solve, and
    zero shifted solutions for the zero columns."""
    G = _graph(2000, 7)
    W = _rhs(G, 8, 9)
    W[:, [1, 6]] = 0.0
    W = torch.from_numpy(W).float().cuda()
    tau = np.array([0.0, 0.1, 1.0, 10.0])
    act = torch.arange(G.n, dtype=torch.int32, device='cuda')
    op = pg.Op(G, precision='fp32')
    kw = dict(track=True, block=True, replace=True, freeze=freeze)
    X, h, it = pg.msolve(op, W, tau, act, 1e-8, **kw)
    _, h2, it2, H = pg.msolve(op, W, tau, act, 1e-8, graph=True, history=512, **kw)
    assert np.array_equal(it, it2) and it[1] == 0 and it[6] == 0 and it.max() > 20
    assert torch.equal(h, h2)
    Xh = H.cols(torch.arange(8, dtype=torch.int32, device='cuda'), act).double().permute(1, 2, 0)
    Xd = X.double()
    assert float(Xh[:, :, [1, 6]].abs().max()) == 0.0
    keep = [0, 2, 3, 4, 5, 7]
    scale = Xd[:, :, keep].abs().amax(dim=1, keepdim=True)
    assert float(((Xh[:, :, keep] - Xd[:, :, keep]).abs() / scale).max()) < 1e-5


@pytest.mark.parametrize('obj', ['er', 'logdet'])
def test_history_rerank_single_pole_objectives(monkeypatch, obj):
    """The single-pole objectives (one pole at 0; logdet without the norm tracker): the history
    re-rank selects as the flushed one, marginals within 1e-6."""
    c = configs.make('c2', 0.02)
    G = pg.Graph(c['n'], c['edges'], c['sigma'])

    def run():
        sk = pg.sketches(G, c['cand'], 16, c['seed'], kind='rademacher', on_device='lazy')
        return pg.polar(G, c['cand'], 5, 0.5, q=16, B=4, tol=c['tol'], sketch=sk, obj=obj, precision='fp32')
    hist, flushed, capped = _greedy_pair(run, monkeypatch)
    _same(hist, flushed)
    _same(capped, flushed)
