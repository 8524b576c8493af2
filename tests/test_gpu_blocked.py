"""Blocked shifted update (POLAR_BLOCK8): X_j, P_j updated once per 8 iterations from the stored
residuals with coefficients built from the same scalars.  The CG itself (alpha, beta, zeta,
freezing, iteration counts, tracked norms, trace terms) is untouched, so those must be equal bit
for bit; X must equal the per-iteration update within rounding (1e-12 relative in FP64), also for
a partial last window, after column compaction, under hipGraph replay and for segments."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from polar_amd import batched, graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _problem(n=3000, R=24, p=12, seed=0):
    G = pg.Graph(n, graphs.chung_lu(n, 2.2, 10.0, 3 + seed), 0.05)
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((n, R)) * rng.uniform(0.1, 10.0, R)
    W[:, 3] = 0.0
    W[:, 7] = 0.0
    W[:, 7][rng.choice(n, 3)] = 1.0                          # a fast column: compaction happens
    tau = pg.fixed_rule(0.5, p, G.sigma, G.Lam(4))[0]
    act = np.sort(rng.choice(n, n // 2, replace=False))
    return G, W, tau, act


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize('graph', [False, True])
def test_blocked_equals_per_iteration_fp64(graph):
    G, W, tau, act = _problem()
    op = pg.Op(G)
    Wd = torch.from_numpy(W).cuda()
    X0, h0, it0, te0 = pg.msolve(op, Wd, tau, act, 1e-10, track=True, trace=True, block=False)
    X1, h1, it1, te1 = pg.msolve(op, Wd, tau, act, 1e-10, track=True, trace=True, block=True, graph=graph)
    assert np.array_equal(it0, it1) and it0[3] == 0
    assert torch.equal(h0, h1) and torch.equal(te0, te1)
    assert _rel(X1, X0) <= 1e-12


def test_blocked_partial_window_at_iteration_cap():
    G, W, tau, act = _problem(R=8)
    op = pg.Op(G)
    Wd = torch.from_numpy(W).cuda()
    outs = []
    for block in (False, True):
        X = torch.empty((len(tau), len(act), 8), dtype=torch.float64, device='cuda')
        with pytest.raises(RuntimeError, match='iteration cap'):
            pg.msolve(op, Wd, tau, act, 1e-14, maxit=13, block=block, out=X)
        outs.append(X)
    assert _rel(outs[1], outs[0]) <= 1e-12


def test_blocked_fp32_close_to_fp64():
    G, W, tau, act = _problem(R=16)
    X64 = pg.msolve(pg.Op(G), torch.from_numpy(W).cuda(), tau, act, 1e-7)[0]
    X32 = pg.msolve(pg.Op(G, precision='fp32'), torch.from_numpy(W).cuda().float(), tau, act, 1e-7, block=True)[0]
    assert _rel(X32.double(), X64) <= 1e-4


def test_blocked_segments_equal_single():
    Gs = [_problem(n, R=8, seed=s)[0] for s, n in enumerate((1500, 3000, 700))]
    U = batched.UnionGraph(Gs)
    rng = np.random.default_rng(5)
    Ws = [rng.standard_normal((G.n, 8)) for G in Gs]
    acts = [np.sort(rng.choice(G.n, G.n // 3, replace=False)) for G in Gs]
    tau = np.stack([pg.fixed_rule(s, 10, G.sigma, G.Lam(4))[0] for s, G in zip((0.25, 0.5, 0.75), Gs)])
    a0 = np.r_[0, np.cumsum([len(a) for a in acts])]
    segs = batched.Segments(*U.segments(a0))
    active = torch.from_numpy(np.concatenate([a + U.row0[g] for g, a in enumerate(acts)]).astype(np.int32)).cuda()
    X, h, it = batched.msolve_batched(pg.Op(U), segs, torch.from_numpy(np.concatenate(Ws)).cuda(), tau, active,
                                      1e-10, track=True, block=True, graph=True)
    for g, G in enumerate(Gs):
        Xs, hs, its = pg.msolve(pg.Op(G), torch.from_numpy(Ws[g]).cuda(), tau[g], acts[g], 1e-10, track=True,
                                block=True)
        assert np.array_equal(it[g], its) and torch.equal(h[:, g], hs)
        assert torch.equal(X[:, a0[g]:a0[g + 1]], Xs)


def test_polar_blocked_fp64_same_greedy():
    G = _problem()[0]
    cand = graphs.candidates(G.n, G.edges, 500, 4)
    poles = pg.fixed_rule(0.5, 12, G.sigma, G.Lam(5))[:2]
    sk = pg.sketches(G, cand, 16, 3, kind='rademacher', on_device='lazy')
    a = pg.polar(G, cand, 5, 0.5, q=16, B=4, sketch=sk, poles=poles)
    b = pg.polar(G, cand, 5, 0.5, q=16, B=4, sketch=sk, poles=poles, block=True)
    assert a['sequence'] == b['sequence'] and a['spmv'] == b['spmv']
    assert np.max(np.abs(np.array(a['marginals']) - b['marginals']) / np.abs(a['marginals'])) <= 1e-10
    assert np.max(np.abs(np.array(a['trace_estimate']) - b['trace_estimate']) / np.abs(a['trace_estimate'])) <= 1e-10


@pytest.mark.parametrize('block', [False, True])
def test_residual_replacement_fp32(block):
    """FP32 vectors with FP64 residual replacement every 24 iterations: the solve still converges,
    and its X is at least as close to the FP64 solution as without replacement (1e-4 bound)."""
    G, W, tau, act = _problem(R=16)
    X64 = pg.msolve(pg.Op(G), torch.from_numpy(W).cuda(), tau, act, 1e-9)[0]
    op32 = pg.Op(G, precision='fp32')
    W32 = torch.from_numpy(W).cuda().float()
    Xp, _, itp = pg.msolve(op32, W32, tau, act, 1e-7, block=block)
    Xr, _, itr = pg.msolve(op32, W32, tau, act, 1e-7, block=block, replace=True)
    ep, er = _rel(Xp.double(), X64), _rel(Xr.double(), X64)
    assert er <= 1e-4 and er <= 2 * ep, (ep, er)
    assert np.all(itr >= 0) and itr.max() < 5 * max(itp.max(), 1)


@pytest.mark.parametrize('precision,block', [('fp64', False), ('fp64', True), ('fp32', True)])
def test_residual_replacement_after_compaction(precision, block):
    """Replacements after an active-column compaction: the even columns are zero, so 8 of 16 slots
    stay active and the solver compacts at its first check; tol 1e-10 keeps it iterating well past
    iteration 24, so every replacement reads the right-hand side through the slot -> column map.
    FP64: X with replacement equals X without it within 1e-8 relative (per-iteration and blocked
    seed iterate).  FP32 (blocked): both within 1e-4 of the FP64 solution, the replaced one no
    worse than twice the plain one or than the FP32 rounding floor (1e-6)."""
    G, W, tau, act = _problem(R=16)
    W[:, 0::2] = 0.0
    W[:, 5] *= 1e-3                              # columns of very different scales
    dt = torch.float64 if precision == 'fp64' else torch.float32
    op = pg.Op(G, precision=precision)
    Wd = torch.from_numpy(W).cuda().to(dt)
    Xp, _, itp = pg.msolve(op, Wd, tau, act, 1e-10, block=block)
    Xr, _, itr = pg.msolve(op, Wd, tau, act, 1e-10, block=block, replace=True)
    assert np.all(itp[0::2] == 0) and np.all(itr[0::2] == 0) and itp.max() > 48
    cols = [c for c in range(1, 16, 2) if np.any(W[:, c])]       # column 3 of the fixture is zero
    assert len(cols) == 7
    if precision == 'fp64':
        for c in cols:
            assert _rel(Xr[:, :, c], Xp[:, :, c]) <= 1e-8, c
    else:
        X64 = pg.msolve(pg.Op(G), torch.from_numpy(W).cuda(), tau, act, 1e-11)[0]
        for c in cols:
            ep, er = _rel(Xp[:, :, c].double(), X64[:, :, c]), _rel(Xr[:, :, c].double(), X64[:, :, c])
            assert er <= 1e-4 and er <= max(2 * ep, 1e-6), (c, ep, er)     # 1e-6: FP32 rounding floor


def test_residual_replacement_fp64_matches_pinned_within_tolerance():
    G, W, tau, act = _problem(R=8)
    op = pg.Op(G)
    Wd = torch.from_numpy(W).cuda()
    X0 = pg.msolve(op, Wd, tau, act, 1e-10)[0]
    X1 = pg.msolve(op, Wd, tau, act, 1e-10, replace=True)[0]
    assert _rel(X1, X0) <= 1e-8


@pytest.mark.parametrize('graph', [False, True])
def test_blocked_column_slices_fp32(graph):
    """FP32 at 72 columns: the window flush runs as column slices of 32, 32 and 8 columns (tiles
    of 32 rows).  CG scalars, tracked norms, trace terms and iteration counts equal the
    per-iteration update bit for bit; X agrees within FP32 rounding."""
    G, W, tau, act = _problem(n=2500, R=72, p=20)
    op = pg.Op(G, precision='fp32')
    Wd = torch.from_numpy(W).cuda().float()
    X0, h0, it0, te0 = pg.msolve(op, Wd, tau, act, 1e-6, track=True, trace=True, block=False)
    X1, h1, it1, te1 = pg.msolve(op, Wd, tau, act, 1e-6, track=True, trace=True, block=True, graph=graph)
    assert np.array_equal(it0, it1)
    assert torch.equal(h0, h1) and torch.equal(te0, te1)
    assert _rel(X1.double(), X0.double()) <= 1e-5


def test_blocked_segments_column_slices_equal_single():
    """Segments whose flush tiles are column slices (FP32, 40 columns: slices of 32 and 8 over the
    plan's segment partition): batched == per-graph bit for bit."""
    Gs = [_problem(n, R=8, seed=s)[0] for s, n in enumerate((1200, 2600, 500))]
    U = batched.UnionGraph(Gs)
    rng = np.random.default_rng(9)
    Ws = [rng.standard_normal((G.n, 40)) for G in Gs]
    acts = [np.sort(rng.choice(G.n, G.n // 2, replace=False)) for G in Gs]
    tau = np.stack([pg.fixed_rule(s, 12, G.sigma, G.Lam(4))[0] for s, G in zip((0.25, 0.5, 0.75), Gs)])
    a0 = np.r_[0, np.cumsum([len(a) for a in acts])]
    segs = batched.Segments(*U.segments(a0))
    active = torch.from_numpy(np.concatenate([a + U.row0[g] for g, a in enumerate(acts)]).astype(np.int32)).cuda()
    W = torch.from_numpy(np.concatenate(Ws)).cuda().float()
    X, h, it = batched.msolve_batched(pg.Op(U, precision='fp32'), segs, W, tau, active, 1e-6, track=True,
                                      block=True)
    for g, G in enumerate(Gs):
        Xs, hs, its = pg.msolve(pg.Op(G, precision='fp32'), torch.from_numpy(Ws[g]).cuda().float(), tau[g],
                                acts[g], 1e-6, track=True, block=True)
        assert np.array_equal(it[g], its) and torch.equal(h[:, g], hs)
        assert torch.equal(X[:, a0[g]:a0[g + 1]], Xs)
