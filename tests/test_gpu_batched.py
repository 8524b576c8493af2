"""Batched form: many graphs as segments of one block-diagonal operator.

Segment g's arithmetic is the single-graph arithmetic, so batched results must equal the
per-graph path exactly (bit for bit, FP64 and FP32): X blocks, tracked norms, trace terms,
iteration counts, and the whole greedy (selections, marginals, trace estimates, counters).
Also covered: hipGraph replay (identical), segments that finish at different iterations
(re-planning), a weighted graph, long rows, added edges (row patch) and uneven shortlists."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import batched, graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _graphs():
    out = []
    e = graphs.barabasi_albert(400, 3, 7)
    out.append(pg.Graph(400, e, 0.1))
    e = graphs.chung_lu(3000, 2.2, 10.0, 3)                      # hub rows longer than one chunk
    out.append(pg.Graph(3000, e, 0.05))
    e = graphs.chung_lu(1500, 2.5, 6.0, 4)
    w = np.random.default_rng(1).uniform(0.5, 1.5, len(e))
    out.append(pg.Graph(1500, e, 0.2, weights=w))
    e = graphs.rmat(2048, 0.57, 0.19, 0.19, 0.05, 8, 5)
    out.append(pg.Graph(2048, e, 0.05))
    return out


def _poles(Gs, p, k=4):
    return [pg.fixed_rule(s, p, G.sigma, G.Lam(k))[:2] for s, G in zip([0.25, 0.5, 0.75, 0.5], Gs)]


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('graph', [False, True])
def test_msolve_batched_equals_single(precision, graph):
    Gs = _graphs()
    S, R, p = len(Gs), 8, 6
    U = batched.UnionGraph(Gs)
    dt = pg._dt(precision)
    rng = np.random.default_rng(11)
    Ws, acts = [], []
    for G in Gs:
        Ws.append(rng.standard_normal((G.n, R)))
        acts.append(np.sort(rng.choice(G.n, G.n // 3, replace=False)))
    Ws[1][:, 5] = 0.0                                             # an inactive column
    poles = _poles(Gs, p)
    tau = np.stack([t for t, _ in poles])
    a0 = np.r_[0, np.cumsum([len(a) for a in acts])]
    segs = batched.Segments(*U.segments(a0))
    active = torch.from_numpy(np.concatenate([a + U.row0[g] for g, a in enumerate(acts)]).astype(np.int32)).cuda()
    W = torch.from_numpy(np.concatenate(Ws)).cuda().to(dt)
    op = pg.Op(U, precision=dt)
    X, h, it, te = batched.msolve_batched(op, segs, W, tau, active, 1e-9, track=True, trace=True, graph=graph)
    assert it[1, 5] == 0 and np.all(it[:, :5] > 0)
    for g, G in enumerate(Gs):
        Xs, hs, its, tes = pg.msolve(pg.Op(G, precision=dt), torch.from_numpy(Ws[g]).cuda().to(dt), tau[g], acts[g],
                                     1e-9, track=True, trace=True)
        assert np.array_equal(it[g], its), g
        assert torch.equal(X[:, a0[g]:a0[g + 1]], Xs), g
        assert torch.equal(h[:, g], hs) and torch.equal(te[:, g], tes), g


def test_msolve_batched_graph_replay_equals_direct():
    """Segments finish at different iterations: the plan shrinks and the captured graphs are rebuilt."""
    Gs = _graphs()
    U = batched.UnionGraph(Gs)
    R, p = 16, 8
    rng = np.random.default_rng(2)
    W = torch.from_numpy(rng.standard_normal((U.n, R))).cuda()
    tau = np.stack([t for t, _ in _poles(Gs, p)])
    acts = [np.arange(G.n) for G in Gs]
    a0 = np.r_[0, np.cumsum([G.n for G in Gs])]
    segs = batched.Segments(*U.segments(a0))
    active = torch.arange(U.n, dtype=torch.int32).cuda()
    op = pg.Op(U)
    res = [batched.msolve_batched(op, segs, W, tau, active, 1e-10, track=True, graph=gr) for gr in (False, True)]
    assert len({tuple(r[2].max(axis=1)) for r in res}) == 1
    assert len(set(res[0][2].max(axis=1))) > 1                   # segments converge at different counts
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])
    del acts


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_polar_batched_equals_per_graph(precision):
    Gs = _graphs()
    k, B, q, p = 5, 3, 16, 8
    cands = [graphs.candidates(G.n, G.edges, 300 + 50 * g, 20 + g) for g, G in enumerate(Gs)]
    cands[0] = cands[0][:6]                             # last round's shortlist is shorter than B (uneven widths)
    poles = _poles(Gs, p, k)
    sk = [pg.sketches(G, cands[g], q, 40 + g, kind='rademacher', on_device=True, precision=precision)
          for g, G in enumerate(Gs)]
    kk = [min(k, len(c)) for c in cands]
    out = batched.polar_batched(Gs, cands, k, 0.5, q=q, B=B, sketch=sk, poles=poles, precision=precision)
    for g, G in enumerate(Gs):
        ref = pg.polar(G, cands[g], kk[g], 0.5, q=q, B=B, sketch=sk[g], poles=poles[g], precision=precision,
                       history=False)
        got = out[g]
        assert got['sequence'][:kk[g]] == ref['sequence'], g
        assert got['marginals'][:kk[g]] == ref['marginals'], g
        assert got['trace_estimate'][:kk[g]] == ref['trace_estimate'][:kk[g]], g
        if kk[g] == k:
            assert (got['spmv'], got['rhs']) == (ref['spmv'], ref['rhs']), g


def test_lazy_sketch_and_split_bank_blocks_change_nothing(monkeypatch):
    """polar() with a DeviceSketch (generated in place) and with separate Z / F bank blocks (the
    memory-tight layout, forced by reporting little free memory) equals the resident-tensor run."""
    G = _graphs()[1]
    cand = graphs.candidates(G.n, G.edges, 400, 3)
    poles = pg.fixed_rule(0.5, 8, G.sigma, G.Lam(4))[:2]
    sk = pg.sketches(G, cand, 16, 9, kind='rademacher', on_device=True)
    ref = pg.polar(G, cand, 4, 0.5, q=16, B=3, sketch=sk, poles=poles)
    lazy = pg.polar(G, cand, 4, 0.5, q=16, B=3, sketch=pg.sketches(G, cand, 16, 9, kind='rademacher', on_device='lazy'),
                    poles=poles)
    total = torch.cuda.mem_get_info()[1]
    monkeypatch.setattr(pg.torch.cuda, 'mem_get_info', lambda *a: (1 << 20, total))
    split = pg.polar(G, cand, 4, 0.5, q=16, B=3, sketch=sk, poles=poles, width=8)
    for out in (lazy, split):
        assert out['sequence'] == ref['sequence'] and out['marginals'] == ref['marginals']
        assert out['trace_estimate'] == ref['trace_estimate'] and out['spmv'] == ref['spmv']


def test_autotune_keeps_results_and_run_batched_cli(tmp_path, monkeypatch):
    """The tuned launch shape changes no result; the c4 CLI (3 graphs, shrunk) agrees across modes."""
    from polar_amd import autotune, run_batched
    monkeypatch.setenv('POLAR_TUNE_CACHE', str(tmp_path / 'tune.json'))
    G = _graphs()[3]
    cand = graphs.candidates(G.n, G.edges, 300, 8)
    poles = pg.fixed_rule(0.5, 8, G.sigma, G.Lam(4))[:2]
    sk = pg.sketches(G, cand, 16, 5, kind='rademacher', on_device='lazy')
    ref = pg.polar(G, cand, 4, 0.5, q=16, B=3, sketch=sk, poles=poles)
    op = pg.Op(G)
    act = torch.from_numpy(np.unique(cand).astype(np.int32)).cuda()
    cfg = autotune.tune(op, 32, poles[0], act, iters=8, reps=1)
    assert set(cfg) >= {'cpl_max', 'grid_max', 'graph'} and (tmp_path / 'tune.json').exists()
    for forced in (cfg, dict(cpl_max=1, grid_max=256, graph=True)):
        out = pg.polar(G, cand, 4, 0.5, q=16, B=3, sketch=sk, poles=poles, tune=forced)
        assert out['sequence'] == ref['sequence'] and out['marginals'] == ref['marginals']
        assert out['spmv'] == ref['spmv']
    res = run_batched.main(['--graphs', '3', '--scale', '0.03', '--k', '3', '--tune'])
    assert res['selections_identical'] and set(res['modes']) == {'batched', 'nograph', 'sequential'}


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_msolve_column_groups_equal_single(precision):
    """Column groups on one graph: each group has its own poles and its own added edges; every
    group's X, tracked norms and iteration counts equal the single solve bit for bit."""
    G = _graphs()[1]
    dt = pg._dt(precision)
    rng = np.random.default_rng(4)
    adds = [[(3, 1700), (10, 2999)], [(3, 5), (400, 401), (10, 11)], []]
    taus = np.stack([pg.fixed_rule(s, 6, G.sigma, G.Lam(4))[0] for s in (0.25, 0.5, 0.75)])
    widths = [8, 4, 4]
    Ws = [rng.standard_normal((G.n, w)) for w in widths]
    act = np.sort(rng.choice(G.n, 900, replace=False))
    opg = pg.Op(G, precision=dt)
    for g, es in enumerate(adds):
        for a, b in es:
            opg.add(a, b, group=g)
    cg = np.concatenate([np.full(w, g) for g, w in enumerate(widths)])
    W = torch.from_numpy(np.concatenate(Ws, axis=1)).cuda().to(dt)
    segs = batched.single_segment(G, len(act))
    X, h, it = batched.msolve_batched(opg, segs, W, taus, torch.from_numpy(act.astype(np.int32)).cuda(), 1e-9,
                                      track=True, col_group=cg, graph=True)
    off = np.r_[0, np.cumsum(widths)]
    for g, es in enumerate(adds):
        op = pg.Op(G, precision=dt)
        for a, b in es:
            op.add(a, b)
        Xs, hs, its = pg.msolve(op, torch.from_numpy(Ws[g]).cuda().to(dt), taus[g], act, 1e-9, track=True)
        assert np.array_equal(it[0, off[g]:off[g + 1]], its), g
        assert torch.equal(X[:, :, off[g]:off[g + 1]], Xs), g
        assert torch.equal(h[:, 0, off[g]:off[g + 1]], hs), g


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_polar_columns_equals_per_s(precision):
    G = _graphs()[1]
    cand = graphs.candidates(G.n, G.edges, 400, 12)
    s_list = [0.25, 0.5, 0.75]
    poles = [pg.fixed_rule(s, 8, G.sigma, G.Lam(5))[:2] for s in s_list]
    sk = pg.sketches(G, cand, 16, 21, kind='rademacher', on_device='lazy')
    out = batched.polar_columns(G, cand, 5, s_list, q=16, B=3, sketch=sk, poles=poles, precision=precision,
                                shared_bank=False)
    out_u = batched.polar_columns(G, cand, 5, s_list, q=16, B=3, sketch=sk, poles=poles, precision=precision,
                                  bank_union=batched.UnionGraph([G] * 3), shared_bank=False)
    for a, b in zip(out, out_u):
        for key in ('sequence', 'marginals', 'trace_estimate', 'spmv', 'rhs'):
            assert a[key] == b[key], key
    for g, s in enumerate(s_list):
        ref = pg.polar(G, cand, 5, s, q=16, B=3, sketch=sk, poles=poles[g], precision=precision)
        got = out[g]
        assert got['sequence'] == ref['sequence'], g
        assert got['marginals'] == ref['marginals'], g
        assert got['trace_estimate'] == ref['trace_estimate'], g
        assert (got['spmv'], got['rhs']) == (ref['spmv'], ref['rhs']), g


@pytest.mark.parametrize('obj', ['er', 'logdet'])
def test_polar_batched_single_pole_objectives_equal_per_graph(obj):
    """The batched greedy with the single-pole objectives (one pole at 0; logdet without the trace
    tracker): one result per graph, equal to polar_gpu.polar run on that graph alone."""
    Gs = _graphs()
    k, B, q = 4, 3, 16
    cands = [graphs.candidates(G.n, G.edges, 200 + 30 * g, 60 + g) for g, G in enumerate(Gs)]
    sk = [pg.sketches(G, cands[g], q, 70 + g, kind='rademacher', on_device=True, precision='fp64')
          for g, G in enumerate(Gs)]
    out = batched.polar_batched(Gs, cands, k, 0.5, q=q, B=B, sketch=sk, obj=obj, precision='fp64')
    for g, G in enumerate(Gs):
        ref = pg.polar(G, cands[g], k, 0.5, q=q, B=B, sketch=sk[g], obj=obj, precision='fp64', history=False)
        assert out[g]['sequence'] == ref['sequence'], g
        assert out[g]['marginals'] == ref['marginals'], g
        assert (out[g]['spmv'], out[g]['rhs'], out[g]['p']) == (ref['spmv'], ref['rhs'], 1), g


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_polar_batched_matches_oracle_per_graph(precision):
    """The batched greedy directly against the oracle (pinned arithmetic) on every graph, with the
    oracle's host Rademacher sketches: FP64 selections and counters exact, marginals and trace
    estimates within 1e-13.  FP32-mixed (residual replacement on): marginals and trace estimates
    within 1e-5, and the same selections up to exact ties -- where the FP32 pick differs, it is in
    the oracle's shortlist with a re-rank score equal to the oracle's best within 1e-6, or both picks
    join two isolated vertices (the R-MAT graph has some; all such pairs share the twin gain);
    rounds are compared up to that round."""
    Gs = _graphs()
    k, B, q, p = 4, 3, 16, 8
    cands = [graphs.candidates(G.n, G.edges, 250 + 40 * g, 80 + g) for g, G in enumerate(Gs)]
    poles = _poles(Gs, p, k)
    Go = [O.Graph(G.n, G.edges, G.sigma, weights=None if G.unit else G.weights) for G in Gs]
    sk = [O.sketches(Go[g], cands[g], q, 90 + g, kind='rademacher') for g in range(len(Gs))]
    out = batched.polar_batched(Gs, cands, k, 0.5, q=q, B=B, sketch=sk, poles=poles, precision=precision)
    bound = 1e-13 if precision == 'fp64' else 1e-5
    for g in range(len(Gs)):
        to = []
        ref = O.polar(Go[g], cands[g], k, 0.5, q=q, B=B, sketch=sk[g], poles=poles[g], width=32, arith='cr',
                      trace=to)
        sg, so = out[g]['sequence'], ref['sequence']
        if precision == 'fp64':
            assert sg == so, g
            assert (out[g]['spmv'], out[g]['rhs']) == (ref['spmv'], ref['rhs']), g
        d = next((r for r, (x, y) in enumerate(zip(sg, so)) if x != y), len(so))
        if d < len(so):
            deg = np.bincount(Gs[g].edges.ravel(), minlength=Gs[g].n)
            used = {int(v) for e in so[:d] for v in cands[g][e]}

            def isolated(e):
                return all(deg[v] == 0 and int(v) not in used for v in cands[g][e])
            sc = np.asarray(to[d]['scores'])
            if sg[d] in to[d]['shortlist']:
                mine = sc[to[d]['shortlist'].index(sg[d])]
                assert abs(mine - sc.max()) <= 1e-6 * abs(sc.max()), (g, d, mine, sc.max())
            else:                               # both picks join two isolated vertices: equal gains
                assert isolated(sg[d]) and isolated(so[d]), (g, d)
        top = min(d + 1, len(so))
        assert _rel(out[g]['marginals'][:top], ref['marginals'][:top]) <= bound, g
        assert _rel(out[g]['trace_estimate'][:top], ref['trace_estimate'][:top]) <= bound, g


@pytest.mark.parametrize('hybrid', [False, True])
def test_polar_columns_matches_oracle_per_s(hybrid):
    """The column-group greedy (three s values on one graph; with hybrid=True the bank runs as row
    segments of a union graph) directly against the oracle run once per s value, FP64."""
    G = _graphs()[1]
    cand = graphs.candidates(G.n, G.edges, 350, 14)
    s_list = [0.25, 0.5, 0.75]
    poles = [pg.fixed_rule(s, 8, G.sigma, G.Lam(4))[:2] for s in s_list]
    Go = O.Graph(G.n, G.edges, G.sigma)
    sk = O.sketches(Go, cand, 16, 23, kind='rademacher')
    out = batched.polar_columns(G, cand, 4, s_list, q=16, B=3, sketch=sk, poles=poles, precision='fp64',
                                bank_union=batched.UnionGraph([G] * 3) if hybrid else None, shared_bank=False)
    for g, s in enumerate(s_list):
        ref = O.polar(Go, cand, 4, s, q=16, B=3, sketch=sk, poles=poles[g], width=32, arith='cr')
        assert out[g]['sequence'] == ref['sequence'], g
        assert (out[g]['spmv'], out[g]['rhs']) == (ref['spmv'], ref['rhs']), g
        assert _rel(out[g]['marginals'], ref['marginals']) <= 1e-13, g
        assert _rel(out[g]['trace_estimate'], ref['trace_estimate']) <= 1e-13, g


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_polar_columns_shared_seed_bank(precision):
    """One seed CG for the bank of every s value (the smallest pole of all is the seed, every group's
    poles its shifted systems): each pole still meets the tolerance, so against the per-s runs of
    the oracle (FP64, pinned arithmetic) the selections are equal, the bank counters are those of a
    single solve, and marginals and trace estimates agree within 1e-5 -- the CG tolerance (1e-9
    here) times the condition number, not rounding.  FP32 is the default user of this path."""
    G = _graphs()[1]
    cand = graphs.candidates(G.n, G.edges, 350, 15)
    s_list = [0.25, 0.5, 0.75]
    poles = [pg.fixed_rule(s, 8, G.sigma, G.Lam(4))[:2] for s in s_list]
    Go = O.Graph(G.n, G.edges, G.sigma)
    sk = O.sketches(Go, cand, 16, 25, kind='rademacher')
    out = batched.polar_columns(G, cand, 4, s_list, q=16, B=3, tol=1e-9, sketch=sk, poles=poles,
                                precision=precision, shared_bank=True)
    assert all(o['shared_bank'] for o in out)
    assert len({o['bank_spmv'] for o in out}) == 1          # one bank solve, counted in every group
    for g, s in enumerate(s_list):
        ref = O.polar(Go, cand, 4, s, q=16, B=3, tol=1e-9, sketch=sk, poles=poles[g], width=32, arith='cr')
        assert out[g]['sequence'] == ref['sequence'], g
        assert out[g]['rhs'] == ref['rhs'], g
        assert _rel(out[g]['marginals'], ref['marginals']) <= 1e-5, g
        assert _rel(out[g]['trace_estimate'], ref['trace_estimate']) <= 1e-5, g
