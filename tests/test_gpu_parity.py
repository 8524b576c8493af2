"""GPU parity of the HIP path against the oracle (oracle/polar_oracle.py) and SciPy.

The oracle runs with arith='cr' (the pinned arithmetic the kernels implement: double-double
reductions, no FMA, fixed operator order), so the FP64 HIP path is compared to it at:
  * integer/index results (sequences, shortlists, iteration counts, spmv/rhs counters): exact;
  * FP64 msolve outputs, marginals, trace estimates: 1e-13 relative (bitwise in practice);
  * FP64 operator products vs SciPy (a different summation order): 1e-12 of the row scale;
  * FP32-mixed marginals and trace estimates vs the FP64 oracle: 1e-5 relative (the bound of the
    FP32-mixed form), with and without FP64 residual replacement.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

torch = pytest.importorskip('torch')

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import configs, graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402
import scalability_inputs as SI  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / 'golden'


def hub_graph(n=3000, seed=5, weighted=False):
    """Power-law graph plus two hubs of degree > 2*256 (exercises the long-row chunk path)."""
    e = graphs.chung_lu(n, 2.3, 8.0, seed)
    hubs = np.array([[0, j] for j in range(1, 900)] + [[7, j] for j in range(300, 1000)])
    e = graphs.simple_edges(np.concatenate([e, hubs]), n)
    w = np.random.default_rng(seed).uniform(0.5, 1.5, len(e)) if weighted else None
    return n, e, w


def dense_patch(n, pairs):
    P = np.zeros((n, n))
    for a, b in pairs:
        P[a, a] += 1
        P[b, b] += 1
        P[a, b] -= 1
        P[b, a] -= 1
    return P


@pytest.mark.parametrize('weighted', [False, True])
@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('R', [1, 3, 8, 16, 64, 128])
def test_spmm_matches_scipy(weighted, precision, R):
    n, e, w = hub_graph(weighted=weighted)
    G = pg.Graph(n, e, 0.3, weights=w)
    assert len(G.schedule()['long_rows']) >= 2
    op = pg.Op(G, precision=precision)
    pairs = [(3, 2500), (0, 2999), (7, 8)]
    for a, b in pairs:
        op.add(a, b)
    rng = np.random.default_rng(R)
    X = rng.standard_normal((n, R))
    M = G.M0.get().toarray() + dense_patch(n, pairs)
    shift = 0.0125
    ref = M @ X + shift * X
    dt = torch.float64 if precision == 'fp64' else torch.float32
    Y = op(torch.from_numpy(X).to('cuda', dt), shift).double().cpu().numpy()
    scale = np.abs(M).sum(1, keepdims=True) * np.abs(X).max()
    tol = 1e-12 if precision == 'fp64' else 2e-6
    assert np.max(np.abs(Y - ref) / scale) < tol


def test_spmm_dropin_no_patch():
    n, e, _ = hub_graph()
    G = pg.Graph(n, e, 1.0)
    X = np.random.default_rng(0).standard_normal((n, 16))
    Y = pg.spmm(G.M0, torch.from_numpy(X).cuda()).cpu().numpy()
    assert np.allclose(Y, G.M0.get() @ X, rtol=0, atol=1e-10)


def _general_csr(n=2500, m=1800, seed=3):
    """A non-Laplacian CSR (n x m, signed random values): empty rows, a row with min(1500, m) entries
    (longer than any chunk), duplicate-free sorted columns."""
    rng = np.random.default_rng(seed)
    A = sp.random(n, m, density=0.004, random_state=seed, format='lil')
    long = min(1500, m)
    A[10, :long] = rng.standard_normal(long)
    A = A.tocsr()
    A.data = rng.uniform(-2.0, 3.0, A.nnz)
    for r in (0, 5, n - 1):
        A.data[A.indptr[r]:A.indptr[r + 1]] = 0.0
    A.eliminate_zeros()
    A.sort_indices()
    assert np.diff(A.indptr)[[0, 5, n - 1]].sum() == 0 and A.indptr[11] - A.indptr[10] >= long
    return A


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('R', [1, 3, 8, 16, 64, 200, 300])
def test_csr_spmm_ctypes_matches_scipy(precision, R):
    """polar_csr_spmm (the replacement of realworld/spmm.py:3-14 csr_spmm) called through ctypes on
    a general CSR, with X and Y at leading dims wider than R: SciPy's A @ X within 1e-12 (FP64) /
    1e-5 (FP32) of the row scale; the columns past R and rows of Y are left untouched."""
    from polar_amd import _lib as L
    lib = L.require_gpu()
    A = _general_csr()
    n, m = A.shape
    dt = torch.float64 if precision == 'fp64' else torch.float32
    rng = np.random.default_rng(R)
    ld = R + 5
    X = torch.from_numpy(rng.standard_normal((m, ld))).to('cuda', dt)
    Y = torch.full((n, ld), 7.0, dtype=dt, device='cuda')
    ip = torch.from_numpy(A.indptr.astype(np.int32)).cuda()
    ix = torch.from_numpy(A.indices.astype(np.int32)).cuda()
    va = torch.from_numpy(A.data).to('cuda', dt)
    rc = lib.polar_csr_spmm(L.dtype_code(dt), n, R, L.ptr(ip), L.ptr(ix), L.ptr(va), L.ptr(X), ld, L.ptr(Y), ld,
                            L.stream_ptr())
    L.check(rc, 'polar_csr_spmm')
    Xh = X.double().cpu().numpy()
    ref = A @ Xh[:, :R]
    Yh = Y.double().cpu().numpy()
    scale = np.maximum(np.abs(A).sum(1).A * np.abs(Xh).max(), 1e-300)
    tol = 1e-12 if precision == 'fp64' else 1e-5
    assert np.max(np.abs(Yh[:, :R] - ref) / scale) < tol
    assert np.all(Yh[[0, 5, n - 1], :R] == 0.0) and np.all(Yh[:, R:] == 7.0)


def test_spmm_accepts_any_csr_and_m0_attributes():
    """spmm(A, X) (realworld/spmm.py:18-25) with a SciPy CSR, a torch sparse CSR tensor and 1-D X
    (the reference's narrow-R path); the drop-in M0 carries indptr / indices / data / shape like the
    reference's cupyx matrix (realworld/spmmscale2.py:4-7), and M0 @ X and spmm(M0, X) agree with
    the general kernel on that CSR."""
    A = _general_csr(n=900, m=900)
    X = np.random.default_rng(1).standard_normal((900, 5))
    Xd = torch.from_numpy(X).cuda()
    ref = A @ X
    assert np.allclose(pg.spmm(A, Xd).cpu().numpy(), ref, rtol=0, atol=1e-11)
    At = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)), torch.from_numpy(A.indices.astype(np.int64)),
                                 torch.from_numpy(A.data), size=A.shape).cuda()
    assert np.allclose(pg.spmm(At, Xd).cpu().numpy(), ref, rtol=0, atol=1e-11)
    y1 = pg.spmm(A, Xd[:, 0]).cpu().numpy()
    assert y1.shape == (900,) and np.allclose(y1, ref[:, 0], rtol=0, atol=1e-11)
    n, e, _ = hub_graph()
    G = pg.Graph(n, e, 0.5)
    M = G.M0.get()
    M0 = G.M0
    assert M0.shape == (n, n) and M0.indptr.dtype == torch.int32 and M0.indices.dtype == torch.int32
    assert M0.data.dtype == torch.float64 and M0.nnz == M.nnz
    assert np.array_equal(M0.indptr.cpu().numpy(), M.indptr) and np.array_equal(M0.indices.cpu().numpy(), M.indices)
    X2 = torch.from_numpy(np.random.default_rng(2).standard_normal((n, 16))).cuda()
    Yg = pg.spmm(M0, X2)
    Yc = pg.csr_spmm(M0, X2)          # the same CSR through the general kernel
    assert torch.allclose(Yg, Yc, rtol=0, atol=1e-10) and torch.allclose(M0 @ X2, Yg, rtol=0, atol=0)
    with pytest.raises(ValueError):
        pg.spmm(A, torch.zeros((7, 2), dtype=torch.float64, device='cuda'))


def _small_problem(seed=0):
    n = 1500
    e = graphs.barabasi_albert(n, 4, seed)
    G0 = O.Graph(n, e, 0.1)
    G = pg.Graph(n, e, 0.1)
    tau, om, _ = pg.fixed_rule(0.5, 12, 0.1, G.Lam(5))
    return n, e, G0, G, tau, om


@pytest.mark.parametrize('freeze', [True, False])
@pytest.mark.parametrize('R', [1, 6, 24])
def test_msolve_matches_oracle(freeze, R):
    n, e, G0, G, tau, om = _small_problem()
    rng = np.random.default_rng(R)
    W = rng.standard_normal((n, R))
    W[:, 0] = 0.0 if R > 1 else W[:, 0]        # an all-zero column stays inactive (polar_gpu.py:94)
    active = np.sort(rng.choice(n, 300, replace=False))
    op0 = O.Op(G0, arith='cr')
    op0.add(5, 900)
    op = pg.Op(G)
    op.add(5, 900)
    Xo, ho, ito, teo = O.msolve(op0, W, tau, active, 1e-9, track=True, freeze=freeze, trace=True)
    Xg, hg, itg, teg = pg.msolve(op, torch.from_numpy(W).cuda(), tau, active, 1e-9, track=True,
                                 freeze=freeze, trace=True)
    assert np.array_equal(ito, itg), (ito, itg)
    Xg = Xg.cpu().numpy()
    assert np.max(np.abs(Xg - Xo)) <= 1e-13 * np.max(np.abs(Xo))
    assert np.max(np.abs(hg.cpu().numpy() - ho)) <= 1e-13 * np.max(np.abs(ho))
    assert np.max(np.abs(teg.cpu().numpy() - teo)) <= 1e-13 * np.max(np.abs(teo))


@pytest.mark.parametrize('R', [1, 20])
def test_msolve_many_poles_few_rows(R):
    """342 poles on 40 active rows: the shifted update spreads the poles over copies of its row
    blocks (pole j to copy j mod ps), with the per-pole scalars staged in LDS (R=1, 342 pole-column
    pairs) and read from global memory (R=20, 6840 pairs). The solutions, h and the iteration
    counts must still equal the oracle's (polar_gpu.py:100-125)."""
    n, e, G0, G, _, _ = _small_problem(seed=3)
    tau, om, _ = pg.fixed_rule(0.5, 342, 0.1, G.Lam(5))
    assert len(tau) == 342
    rng = np.random.default_rng(40 + R)
    W = rng.standard_normal((n, R))
    active = np.sort(rng.choice(n, 40, replace=False))
    op0 = O.Op(G0, arith='cr')
    Xo, ho, ito = O.msolve(op0, W, tau, active, 1e-10, track=True, freeze=True)
    Xg, hg, itg = pg.msolve(pg.Op(G), torch.from_numpy(W).cuda(), tau, active, 1e-10, track=True,
                            freeze=True)
    assert np.array_equal(ito, itg), (ito, itg)
    Xg = Xg.cpu().numpy()
    assert Xg.shape == Xo.shape == (len(tau), len(active), R)
    assert np.max(np.abs(Xg - Xo)) <= 1e-13 * np.max(np.abs(Xo))
    assert np.max(np.abs(hg.cpu().numpy() - ho)) <= 1e-13 * np.max(np.abs(ho))


def test_msolve_iteration_cap():
    n, e, G0, G, tau, om = _small_problem()
    W = np.random.default_rng(0).standard_normal((n, 4))
    with pytest.raises(RuntimeError, match='multi-shift CG iteration cap'):
        pg.msolve(pg.Op(G), torch.from_numpy(W).cuda(), tau, np.arange(10), 1e-12, maxit=16)


def test_select_topb_ties_and_dead():
    from polar_amd import _lib as L
    lib = L.require_gpu()
    rng = np.random.default_rng(1)
    s = np.round(rng.standard_normal(100_003), 1)
    s[rng.choice(len(s), 500, replace=False)] = -np.inf
    sd = torch.from_numpy(s).cuda()
    for Br in (1, 5, 16, 64):
        out = torch.empty(64, dtype=torch.int32, device='cuda')
        host = np.zeros(Br, dtype=np.int32)
        ws = torch.empty(lib.polar_select_ws_bytes(len(s)), dtype=torch.uint8, device='cuda')
        L.check(lib.polar_select_topb(L.ptr(sd), len(s), Br, L.ptr(out), host.ctypes.data_as(L.vp), L.ptr(ws),
                                      ws.numel(), L.stream_ptr()))
        assert host.tolist() == np.argsort(-s, kind='stable')[:Br].tolist()


def test_rademacher_device_matches_host():
    n, e, w = hub_graph(n=2000, weighted=True)
    G = pg.Graph(n, e, 0.1, weights=w)
    cand = graphs.candidates(n, e, 300, 3)
    Zh, Fh, Gh = pg.sketches(G, cand, 24, 11, kind='rademacher')
    Zd, Fd, Gd = pg.sketches(G, cand, 24, 11, kind='rademacher', on_device=True)
    assert np.array_equal(Zd.cpu().numpy(), Zh)
    assert np.max(np.abs(Fd.cpu().numpy() - Fh)) < 1e-12
    assert np.array_equal(np.stack([Gd[i] for i in range(len(cand))]), Gh)
    Zo, Fo, Go = O.sketches(O.Graph(n, e, 0.1, weights=w), cand, 24, 11, kind='rademacher')
    assert np.array_equal(Zo, Zh) and np.array_equal(Go, Gh) and np.max(np.abs(Fo - Fh)) < 1e-12


def _c1_inputs(scale=1.0):
    c = configs.make('c1', scale)
    Go = O.Graph(c['n'], c['edges'], c['sigma'])
    sk = O.sketches(Go, c['cand'], c['q'], c['seed'], kind='rademacher')
    return c, Go, sk


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_polar_c1_fp64_parity():
    """c1 (BA n=2000, q=32, 16 poles, N=5000, k=10, B=1): identical sequence, identical spmv / rhs
    counters and per-solve column-iterations; marginals and trace estimates within 1e-13."""
    c, Go, sk = _c1_inputs()
    kw = dict(q=c['q'], B=c['B'], tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']))
    ro = O.polar(Go, c['cand'], c['k'], c['s'], width=64, arith='cr', **kw)
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    rg = pg.polar(G, c['cand'], c['k'], c['s'], **kw)
    assert rg['sequence'] == ro['sequence']
    assert rg['rhs'] == ro['rhs'] == 2 * c['q'] + c['k'] * c['B']
    assert rg['spmv'] == ro['spmv']
    assert _rel(rg['marginals'], ro['marginals']) < 1e-13
    assert _rel(rg['trace_estimate'], ro['trace_estimate']) < 1e-13


def test_polar_b4_and_naive_parity():
    c = configs.make('c2', 0.02)
    Go = O.Graph(c['n'], c['edges'], c['sigma'])
    sk = O.sketches(Go, c['cand'], 16, c['seed'], kind='rademacher')
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    kw = dict(q=16, tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']))
    for B, naive in ((4, False), (1, False), (1, True)):
        ro = O.polar(Go, c['cand'], 6, c['s'], B=B, naive=naive, width=32, arith='cr', **kw)
        rg = pg.polar(G, c['cand'], 6, c['s'], B=B, naive=naive, **kw)
        assert rg['sequence'] == ro['sequence'], (B, naive)
        assert rg['spmv'] == ro['spmv'], (B, naive)
        assert _rel(rg['marginals'], ro['marginals']) < 1e-13
        assert _rel(rg['trace_estimate'], ro['trace_estimate']) < 1e-13


def test_polar_tie_gap_records():
    """Per-round tie records: the shortlist's proxy scores and re-rank scores (trace entries) equal
    the oracle's, the winner is the first argmax of the re-rank scores, and proxy_gap / rerank_gap
    are the relative top-2 gaps of those scores."""
    c = configs.make('c2', 0.02)
    Go = O.Graph(c['n'], c['edges'], c['sigma'])
    sk = O.sketches(Go, c['cand'], 16, c['seed'], kind='rademacher')
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    kw = dict(q=16, B=4, tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']))
    to, tg = [], []
    ro = O.polar(Go, c['cand'], 6, c['s'], width=32, arith='cr', trace=to, **kw)
    rg = pg.polar(G, c['cand'], 6, c['s'], trace=tg, **kw)
    assert rg['sequence'] == ro['sequence'] and len(rg['proxy_gap']) == len(rg['rerank_gap']) == 6
    for r, (a, b) in enumerate(zip(tg, to)):
        assert a['shortlist'] == b['shortlist'] and a['chosen'] == b['chosen']
        assert _rel(a['proxy'], b['proxy']) < 1e-13 and _rel(a['scores'], b['scores']) < 1e-13
        sc = np.asarray(a['scores'])
        assert a['chosen'] == a['shortlist'][int(np.argmax(sc))]
        top = np.sort(sc)[::-1]
        assert abs(rg['rerank_gap'][r] - (top[0] - top[1]) / abs(top[0])) <= 1e-12
        assert abs(rg['proxy_gap'][r] - (a['proxy'][0] - a['proxy'][1]) / abs(a['proxy'][0])) <= 1e-12
        assert rg['proxy_gap'][r] >= 0 and rg['rerank_gap'][r] >= 0


@pytest.mark.parametrize('idx', [0, 1, 5])
def test_reference_fixture_sequences(idx):
    """synthetic/scalability/results.json sequences (Naive-ShiftedCG and Inc-Full) through the HIP
    path: single-RHS variant = no pole freezing, quad rule p=20, Gaussian q=24 sketches."""
    fx = json.loads((GOLD / 'scalability_fixture.json').read_text())['configurations'][idx]
    d = SI.make(fx['n'], fx['topology'], fx['support'])
    tau, om = SI.quad(0.5, 20)
    G = pg.Graph(d['n'], d['edges'], 1.0)
    sk = (d['Z'], d['F'], d['G'])
    for naive, name in ((True, 'Naive-ShiftedCG'), (False, 'Inc-Full')):
        out = pg.polar(G, d['cand'], 15, 0.5, B=1, tol=1e-7, sketch=sk, naive=naive, poles=(tau, om),
                       width=48, freeze=False)
        assert out['sequence'] == fx['methods'][name]['sequence'], name
        sp_ = out['spmv'] if naive else sum(out['solves'][:-1])
        assert abs(sp_ - fx['methods'][name]['spmvs']) <= 0.01 * fx['methods'][name]['spmvs']


@pytest.mark.parametrize('replace', [True, False])
def test_fp32_mixed_vs_fp64(replace):
    """FP32 vectors / FP64 recurrences (residual replacement on, the default, and off): same c1
    sequence; marginals and trace estimate within 1e-5 relative of the FP64 oracle."""
    c, Go, sk = _c1_inputs()
    kw = dict(q=c['q'], B=c['B'], tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']))
    ro = O.polar(Go, c['cand'], c['k'], c['s'], width=64, arith='cr', **kw)
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    rg = pg.polar(G, c['cand'], c['k'], c['s'], precision='fp32', replace=replace, **kw)
    assert rg['replace'] is replace
    assert rg['sequence'] == ro['sequence']
    assert _rel(rg['marginals'], ro['marginals']) < 1e-5
    assert _rel(rg['trace_estimate'], ro['trace_estimate']) < 1e-5


def test_true_marginals_vs_eigenvalues():
    n = 600
    e = graphs.barabasi_albert(n, 3, 4)
    G = pg.Graph(n, e, 0.5)
    cand = graphs.candidates(n, e, 200, 9)
    seq = pg.random_set(cand, 5, 3)
    g, cert = pg.true_marginals(G, cand, seq, 0.5)
    exact = np.diff(np.r_[0.0, O.eval_prefix(G.M0.get().toarray(), cand, seq, 0.5)])
    assert np.max(np.abs(g - exact) / exact) < 1e-7
    g_o, cert_o = O.true_marginals(O.Graph(n, e, 0.5), cand, seq, 0.5, arith='cr')
    assert _rel(g, g_o) < 1e-13
    assert cert['p'] == cert_o['p'] and abs(cert['residual_rho'] - cert_o['residual_rho']) < 1e-20


@pytest.mark.parametrize('obj,precision', [('frac', 'fp64'), ('logdet', 'fp64'), ('frac', 'fp32')])
def test_true_marginals_positions_together(obj, precision):
    """The positions of the sequence as column groups of one solve (each sees the edges before it)
    equal the sequential loop: same values and spmv, in chunks that do not divide the length (FP32
    with the blocked update: the same CG scalars, values within its rounding)."""
    n = 700
    e = graphs.barabasi_albert(n, 3, 6)
    G = pg.Graph(n, e, 0.4)
    cand = graphs.candidates(n, e, 300, 5)
    seq = pg.random_set(cand, 11, 2)
    kw = dict(obj=obj, precision=precision, tol=1e-10 if precision == 'fp64' else 1e-6)
    ref, c_ref = pg.true_marginals(G, cand, seq, 0.5, together=1, **kw)
    for together in (4, 64):
        g, c = pg.true_marginals(G, cand, seq, 0.5, together=together, **kw)
        assert c['spmv'] == c_ref['spmv'] and c['p'] == c_ref['p']
        assert _rel(g, ref) < (1e-13 if precision == 'fp64' else 1e-5), together


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_msolve_column_compaction(precision):
    """Half the columns are zero (inactive from the start) and the rest mix dense random columns
    with e_a - e_b pairs (fast convergence): the solver compacts its active columns several times.
    FP64 must still equal the oracle bit for bit (columns are independent); FP32 within 1e-4."""
    n, e, G0, G, tau, om = _small_problem(seed=2)
    R = 32
    rng = np.random.default_rng(7)
    W = np.zeros((n, R))
    for c in range(1, R, 2):
        if c % 4 == 1:
            W[:, c] = rng.standard_normal(n)
        else:
            a, b = rng.choice(n, 2, replace=False)
            W[a, c], W[b, c] = 1.0, -1.0
    active = np.sort(rng.choice(n, 200, replace=False))
    op0 = O.Op(G0, arith='cr')
    Xo, ho, ito, teo = O.msolve(op0, W, tau, active, 1e-10, track=True, trace=True)
    dt = torch.float64 if precision == 'fp64' else torch.float32
    Xg, hg, itg, teg = pg.msolve(pg.Op(G, precision=precision), torch.from_numpy(W).to('cuda', dt), tau, active,
                                 1e-10, track=True, trace=True)
    Xg = Xg.double().cpu().numpy()
    if precision == 'fp64':
        assert np.array_equal(ito, itg), (ito, itg)
        assert np.max(np.abs(Xg - Xo)) <= 1e-13 * np.max(np.abs(Xo))
        assert np.max(np.abs(hg.cpu().numpy() - ho)) <= 1e-13 * np.max(np.abs(ho))
        assert np.max(np.abs(teg.cpu().numpy() - teo)) <= 1e-13 * np.max(np.abs(teo))
    else:
        assert np.all(itg[0::2] == 0)
        assert np.max(np.abs(Xg - Xo)) <= 1e-4 * np.max(np.abs(Xo))
        assert np.max(np.abs(hg.cpu().numpy() - ho) / np.maximum(ho, 1e-300)) <= 1e-4


def test_slq_greedy_matches_oracle():
    """SLQ-Greedy baseline (realworld/slq.py): same probes, same recurrences; the GPU estimates
    equal the oracle's to 1e-9 relative and the selections are identical."""
    from polar_amd import slq
    n = 300
    e = graphs.barabasi_albert(n, 3, 8)
    cand = graphs.candidates(n, e, 40, 9)
    Gh = pg.Graph(n, e, 0.2)
    Go = O.Graph(n, e, 0.2)
    got = slq.slq_greedy(Gh, cand, 3, 0.5, probes=6, m=14, chunk=60, seed=11, return_estimates=True)
    ref = O.slq_greedy(Go, cand, 3, 0.5, probes=6, m=14, chunk=60, seed=11)
    assert got['sequence'] == ref['sequence'] and got['spmv'] == ref['spmv']
    for a, b in zip(got['estimates'], ref['estimates']):
        assert np.max(np.abs(a - b) / np.abs(b)) <= 1e-9


def test_polar_many_poles_fp64_parity():
    """150 poles: the h/t maintenance runs in LDS tiles of 12 pairs (the last one partial) and
    the rank-1 pass over 150 poles; the greedy equals the pinned oracle (sequence, counters,
    marginals and traces within 1e-13)."""
    c = configs.make('c2', 0.015)
    tau, om = pg.fixed_rule(0.5, 150, c['sigma'], pg.Graph(c['n'], c['edges'], c['sigma']).Lam(5))[:2]
    Go = O.Graph(c['n'], c['edges'], c['sigma'])
    cand = c['cand'][:503]
    sk = O.sketches(Go, cand, 8, c['seed'], kind='rademacher')
    kw = dict(q=8, tol=1e-7, sketch=sk, poles=(tau, om))
    ro = O.polar(Go, cand, 4, 0.5, B=3, width=16, arith='cr', **kw)
    rg = pg.polar(pg.Graph(c['n'], c['edges'], c['sigma']), cand, 4, 0.5, B=3, **kw)
    assert rg['sequence'] == ro['sequence']
    assert rg['spmv'] == ro['spmv'] and rg['p'] == 150
    assert _rel(rg['marginals'], ro['marginals']) < 1e-13
    assert _rel(rg['trace_estimate'], ro['trace_estimate']) < 1e-13


@pytest.mark.parametrize('obj', ['er', 'logdet'])
def test_polar_er_logdet_parity(obj):
    """The single-pole objectives through the HIP path (proxy obj code, log1p re-rank for logdet,
    no trace tracker for logdet): the greedy and the certified marginals equal the oracle's."""
    n = 400
    e = graphs.barabasi_albert(n, 3, 13)
    cand = graphs.candidates(n, e, 300, 8)
    Go, G = O.Graph(n, e, 0.3), pg.Graph(n, e, 0.3)
    sk = O.sketches(Go, cand, 16, 4, kind='rademacher')
    kw = dict(q=16, B=4, tol=1e-9, obj=obj, sketch=sk)
    ro = O.polar(Go, cand, 5, 0.5, arith='cr', **kw)
    rg = pg.polar(G, cand, 5, 0.5, **kw)
    assert rg['sequence'] == ro['sequence']
    assert (rg['spmv'], rg['rhs'], rg['p']) == (ro['spmv'], ro['rhs'], 1)
    assert _rel(rg['marginals'], ro['marginals']) < 1e-13
    if obj == 'er':
        assert _rel(rg['trace_estimate'], ro['trace_estimate']) < 1e-13
    g_o, c_o = O.true_marginals(Go, cand, ro['sequence'], 0.5, obj=obj, arith='cr')
    g_g, c_g = pg.true_marginals(G, cand, ro['sequence'], 0.5, obj=obj)
    assert _rel(g_g, g_o) < 1e-12 and c_g['spmv'] == c_o['spmv'] and c_g['p'] == 1


@pytest.mark.parametrize('kind', ['block', 'gauss'])
def test_polar_reference_sketch_kinds_parity(kind):
    """polar() generating its own sketches of the reference's default kinds (block: stacked
    orthogonal rows; gauss) from the seed: same sequence, counters and marginals as the oracle."""
    n = 500
    e = graphs.barabasi_albert(n, 3, 17)
    cand = graphs.candidates(n, e, 200, 9)
    Go, G = O.Graph(n, e, 0.2), pg.Graph(n, e, 0.2)
    tau, om, _ = pg.fixed_rule(0.5, 8, 0.2, G.Lam(4))
    kw = dict(q=16, B=3, tol=1e-9, seed=31, kind=kind, poles=(tau, om))
    ro = O.polar(Go, cand, 4, 0.5, arith='cr', **kw)
    rg = pg.polar(G, cand, 4, 0.5, **kw)
    assert rg['sequence'] == ro['sequence'] and (rg['spmv'], rg['rhs']) == (ro['spmv'], ro['rhs'])
    assert _rel(rg['marginals'], ro['marginals']) < 1e-13
    assert _rel(rg['trace_estimate'], ro['trace_estimate']) < 1e-13


def test_polar_time_limit_truncates_after_first_round():
    """time_limit (polar_gpu.py:215): the limit is checked before every round after the first, so
    a zero limit returns one round, truncated, with the oracle's first pick and counters."""
    n = 300
    e = graphs.barabasi_albert(n, 3, 5)
    cand = graphs.candidates(n, e, 100, 3)
    Go, G = O.Graph(n, e, 0.2), pg.Graph(n, e, 0.2)
    tau, om, _ = pg.fixed_rule(0.5, 6, 0.2, G.Lam(4))
    sk = O.sketches(Go, cand, 8, 2, kind='rademacher')
    kw = dict(q=8, B=2, tol=1e-9, sketch=sk, poles=(tau, om), time_limit=0.0)
    ro = O.polar(Go, cand, 4, 0.5, arith='cr', **kw)
    rg = pg.polar(G, cand, 4, 0.5, **kw)
    assert rg['rounds'] == ro['rounds'] == 1 and rg['truncated'] and ro['truncated']
    assert rg['sequence'] == ro['sequence'] and (rg['spmv'], rg['rhs']) == (ro['spmv'], ro['rhs'])


@pytest.mark.parametrize('obj', ['er', 'logdet'])
def test_fp32_mixed_single_pole_objectives(obj):
    """FP32-mixed with the single-pole objectives against the FP64 HIP path: same selections on a
    graph with distinct gains; marginals within 1e-5 relative."""
    n = 400
    e = graphs.barabasi_albert(n, 3, 13)
    cand = graphs.candidates(n, e, 300, 8)
    G = pg.Graph(n, e, 0.3)
    sk = O.sketches(O.Graph(n, e, 0.3), cand, 16, 4, kind='rademacher')
    kw = dict(q=16, B=4, tol=1e-9, obj=obj, sketch=sk)
    r64 = pg.polar(G, cand, 5, 0.5, **kw)
    r32 = pg.polar(G, cand, 5, 0.5, precision='fp32', **kw)
    assert r32['sequence'] == r64['sequence']
    assert _rel(r32['marginals'], r64['marginals']) < 1e-5


def test_polar_weighted_hub_graph_parity():
    """A weighted graph with two hubs of degree > 700 (weighted operator kernels, long-row chunks
    and their finisher inside every solve, added unit-weight edges on top): the FP64 greedy equals
    the oracle; FP32-mixed keeps the selections."""
    n, e, w = hub_graph(n=1500, seed=9, weighted=True)
    cand = graphs.candidates(n, e, 250, 4)
    Go, G = O.Graph(n, e, 0.2, weights=w), pg.Graph(n, e, 0.2, weights=w)
    tau, om, _ = pg.fixed_rule(0.5, 8, 0.2, G.Lam(4))
    sk = O.sketches(Go, cand, 8, 6, kind='rademacher')
    kw = dict(q=8, B=3, tol=1e-9, sketch=sk, poles=(tau, om))
    ro = O.polar(Go, cand, 4, 0.5, arith='cr', **kw)
    rg = pg.polar(G, cand, 4, 0.5, **kw)
    assert rg['sequence'] == ro['sequence'] and (rg['spmv'], rg['rhs']) == (ro['spmv'], ro['rhs'])
    assert _rel(rg['marginals'], ro['marginals']) < 1e-13
    assert _rel(rg['trace_estimate'], ro['trace_estimate']) < 1e-13
    r32 = pg.polar(G, cand, 4, 0.5, precision='fp32', **kw)
    assert r32['replace'] is True
    assert r32['sequence'] == ro['sequence'] and _rel(r32['marginals'], ro['marginals']) < 1e-5
