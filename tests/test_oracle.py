"""Pins of the CPU oracle (oracle/polar_oracle.py) against the reference's fixtures and against
closed forms.  Each check is chosen so that a plausible mistake (a dropped term, a wrong shift or
index, an off-by-one in freezing or convergence) fails it.  CPU only."""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.special as ss

from oracle import polar_oracle as O
from polar_amd import configs, graphs, quadrature
import scalability_inputs as SI

GOLD = Path(__file__).resolve().parent / 'golden'


# ---------------------------------------------------------------- quadrature
def test_composite_pole_counts_match_reference_runs():
    """p=(K+2)r of every timed real graph (realworld/results/time_*.json: dmax, sigma=1, k=20,
    eps=1e-2) -- pins r, K and Lam(k)=sigma+2*dmax+4k."""
    meta = json.loads((GOLD / 'realworld_meta.json').read_text())
    for row in meta['time']:
        lam = row['sigma'] + 2 * row['dmax'] + 4 * row['k']
        for f in (O.composite, quadrature.composite):
            tau, om, _ = f(row['s'], row['sigma'], lam, row['eps'])
            assert len(tau) == row['p']['POLAR(B=1)'], row['graph']


def test_certified_evaluator_constants_match_reference():
    """effect_com-Youtube.json `_evaluator`: p=342, quadrature_bound, residual_rho (which pins the
    smallest pole of the eps=1e-9 rule and true_marginals' rho formula)."""
    meta = json.loads((GOLD / 'realworld_meta.json').read_text())
    for ev in meta['evaluator']:
        lam = ev['sigma'] + 2 * ev['dmax'] + 4 * ev['k']
        obj = 'frac' if ev['target'].startswith('frac') else ev['target']
        tau, om, bound = O.poles(obj, 0.5, ev['sigma'], lam, 1e-9)
        assert len(tau) == ev['p']
        assert bound == pytest.approx(ev['quadrature_bound'], rel=0, abs=1e-25)
        rho = 1e-10 * float(np.max((lam + tau) / (ev['sigma'] + tau)))
        assert rho == pytest.approx(ev['residual_rho'], rel=1e-12)


@pytest.mark.parametrize('s', [0.25, 0.5, 0.75])
def test_composite_closed_form(s):
    """sum_j w_j/(x+t_j) approximates x^-s within 256*4^-r on [a, z] (verify_stieltjes.py:43-52)."""
    a, z = 0.3, 900.0
    for eps in (1e-2, 1e-5):
        tau, om, bound = O.composite(s, a, z, eps)
        x = np.geomspace(a, z, 400)
        approx = (om[None] / (x[:, None] + tau[None])).sum(1)
        assert np.max(np.abs(approx * x ** s - 1)) <= bound
        assert np.all(np.diff(tau) > 0) and np.all(om > 0)
        t2, o2, _ = quadrature.composite(s, a, z, eps)
        assert np.allclose(t2, tau, rtol=1e-14) and np.allclose(o2, om, rtol=1e-13)


def test_jacobi_panel_matches_scipy():
    """The u^beta panel equals SciPy's Gauss-Jacobi rule mapped to [0,1]."""
    for beta in (-0.5, -0.25, 0.3):
        x, w = O._jacobi_rule(beta, 9)
        y, v = ss.roots_jacobi(9, 0.0, beta)
        assert np.allclose(np.sort(x), (1 + np.sort(y)) / 2, atol=1e-14)
        assert np.allclose(w[np.argsort(x)], v[np.argsort(y)] / 2 ** (beta + 1), rtol=1e-12)


def test_fixed_rule_is_the_scalability_rule_rescaled():
    t0, w0 = SI.quad(0.5, 20)
    t1, w1, _ = O.fixed_rule(0.5, 20, 1.0, 1.0)
    assert np.allclose(t0, t1, rtol=1e-14) and np.allclose(w0, w1, rtol=1e-13)
    tau, om, _ = quadrature.fixed_rule(0.5, 16, 0.1, 340.0)
    x = np.geomspace(0.1, 340, 500)
    assert np.max(np.abs((om[None] / (x[:, None] + tau[None])).sum(1) * x ** 0.5 - 1)) < 1e-3


# ---------------------------------------------------------------- operator and solver
def _problem(n=300, seed=3):
    e = graphs.barabasi_albert(n, 3, seed)
    w = np.random.default_rng(seed).uniform(0.5, 1.5, len(e))
    return n, e, w


def test_graph_and_op_against_dense():
    n, e, w = _problem()
    G = O.Graph(n, e, 0.2, weights=w)
    M = np.zeros((n, n))
    for (a, b), x in zip(e, w):
        M[a, a] += x
        M[b, b] += x
        M[a, b] -= x
        M[b, a] -= x
    M += 0.2 * np.eye(n)
    assert np.allclose(G.M0.toarray(), M)
    op = O.Op(G)
    op.add(1, 250)
    op.add(4, 7)
    Madd = M.copy()
    for a, b in ((1, 250), (4, 7)):
        Madd[a, a] += 1
        Madd[b, b] += 1
        Madd[a, b] -= 1
        Madd[b, a] -= 1
    X = np.random.default_rng(0).standard_normal((n, 5))
    assert np.allclose(op(X, 0.7), Madd @ X + 0.7 * X)


@pytest.mark.parametrize('freeze', [True, False])
def test_msolve_against_dense_solves(freeze):
    """validate.py:14-24: shifted solutions on the active rows, tracked ||u_j||^2, and the trace
    tracker w^T (M + tau_j)^-1 w, against numpy.linalg.solve."""
    n, e, w = _problem()
    G = O.Graph(n, e, 0.2, weights=w)
    op = O.Op(G)
    op.add(0, 299)
    M = G.M0.toarray()
    M[0, 0] += 1
    M[299, 299] += 1
    M[0, 299] -= 1
    M[299, 0] -= 1
    tau, om, _ = O.composite(0.5, 0.2, G.Lam(2), 1e-2)
    rng = np.random.default_rng(1)
    W = rng.standard_normal((n, 4))
    W[:, 2] = 0.0
    active = np.sort(rng.choice(n, 40, replace=False))
    X, h, it, te = O.msolve(op, W, tau, active, 1e-12, track=True, freeze=freeze, trace=True)
    assert it[2] == 0 and np.all(X[:, :, 2] == 0) and np.all(it[[0, 1, 3]] > 0)
    for j in range(0, len(tau), 7):
        U = np.linalg.solve(M + tau[j] * np.eye(n), W)
        assert np.max(np.abs(U[active] - X[j])) < 1e-9 * np.abs(U).max()
        assert np.allclose(h[j], np.sum(U * U, 0), rtol=1e-8)
        assert np.allclose(te[j], np.sum(W * U, 0), rtol=1e-8)


def test_msolve_iteration_cap_and_zero_rhs():
    n, e, w = _problem()
    G = O.Graph(n, e, 0.2)
    tau, om, _ = O.composite(0.5, 0.2, G.Lam(1), 1e-2)
    W = np.random.default_rng(2).standard_normal((n, 2))
    with pytest.raises(RuntimeError, match='iteration cap'):
        O.msolve(O.Op(G), W, tau, np.arange(5), 1e-14, maxit=7)
    X, h, it = O.msolve(O.Op(G), np.zeros((n, 3)), tau, np.arange(5), 1e-7, track=True)
    assert np.all(it == 0) and np.all(X == 0) and np.all(h == 0)


# ---------------------------------------------------------------- greedy: reference fixture
FIX = json.loads((GOLD / 'scalability_fixture.json').read_text())


@pytest.mark.parametrize('idx', [0, 1, 2])
def test_polar_reproduces_reference_fixture(idx):
    """Naive-ShiftedCG and Inc-Full sequences and spmv counts of the reference's CPU fixture
    (synthetic/scalability/results.json): regenerated seeded inputs, quad rule p=20, q=24 Gaussian
    sketches, single-RHS variant without pole freezing.  Inc-Full does k-1 maintenance solves,
    POLAR(B=1) k re-rank solves: compare bank + first k-1 solves."""
    fx = FIX['configurations'][idx]
    d = SI.make(fx['n'], fx['topology'], fx['support'])
    assert len(d['edges']) == fx['m'] and len(np.unique(d['cand'])) == fx['n_candidates_vertices']
    tau, om = SI.quad(0.5, 20)
    G = O.Graph(d['n'], d['edges'], 1.0)
    sk = (d['Z'], d['F'], d['G'])
    for naive, name in ((True, 'Naive-ShiftedCG'), (False, 'Inc-Full')):
        out = O.polar(G, d['cand'], 15, 0.5, B=1, tol=1e-7, sketch=sk, naive=naive, poles=(tau, om),
                      width=48, freeze=False)
        ref = fx['methods'][name]
        assert out['sequence'] == ref['sequence'], name
        spmv = out['spmv'] if naive else sum(out['solves'][:-1])
        assert spmv == ref['spmvs'], name


def test_polar_b1_equals_naive_and_counters():
    """validate.py:45-52 / every time_*.json: POLAR(B=1) selects the Naive sequence; rhs = 2q+kB
    for POLAR and k*2q for Naive (with pole freezing on, as polar_gpu.py runs)."""
    n = 400
    e = graphs.barabasi_albert(n, 3, 1)
    G = O.Graph(n, e, 0.1)
    cand = graphs.candidates(n, e, 600, 2)
    sk = O.sketches(G, cand, 12, 5, kind='rademacher')
    tau, om, _ = O.fixed_rule(0.5, 12, 0.1, G.Lam(6))
    kw = dict(q=12, tol=1e-8, sketch=sk, poles=(tau, om))
    nv = O.polar(G, cand, 6, 0.5, naive=True, **kw)
    b1 = O.polar(G, cand, 6, 0.5, B=1, **kw)
    b4 = O.polar(G, cand, 6, 0.5, B=4, **kw)
    assert nv['sequence'] == b1['sequence']
    assert b1['rhs'] == 2 * 12 + 6 * 1 and b4['rhs'] == 2 * 12 + 6 * 4 and nv['rhs'] == 6 * 2 * 12
    assert len(set(b4['sequence'])) == 6


def test_polar_matches_dense_closed_form_greedy():
    """Dense closed form of the same algorithm: X_j=(M+tau_j)^-1 Z, Y_j=(M+tau_j)^-1 F exactly,
    proxy top-B, exact re-rank marginal, sketch updates by re-solving; also the per-round
    marginal G_q and the sketched trace sum_j om_j tr(Z^T (M_S+tau_j)^-1 Z)."""
    n = 250
    e = graphs.barabasi_albert(n, 3, 7)
    sigma = 0.1
    G = O.Graph(n, e, sigma)
    cand = graphs.candidates(n, e, 400, 8)
    q = 10
    Z, F, Gc = O.sketches(G, cand, q, 3, kind='rademacher')
    tau, om, _ = O.fixed_rule(0.5, 10, sigma, G.Lam(5))
    k, B = 5, 3
    out = O.polar(G, cand, k, 0.5, q=q, B=B, tol=1e-12, sketch=(Z, F, Gc), poles=(tau, om))
    M = G.M0.toarray()
    Fc = F.copy()
    live = np.ones(len(cand), bool)
    seq, marg, tr = [], [], []
    for r in range(k):
        inv = [np.linalg.inv(M + t * np.eye(n)) for t in tau]
        tr.append(sum(w * np.trace(Z.T @ A @ Z) for w, A in zip(om, inv)))
        X = [A @ Z for A in inv]
        Y = [A @ Fc for A in inv]
        h = np.array([np.sum((Xj[cand[:, 0]] - Xj[cand[:, 1]]) ** 2, 1) for Xj in X])
        t = np.array([np.sum((Yj[cand[:, 0]] - Yj[cand[:, 1]]) ** 2, 1) for Yj in Y])
        sc = np.sum(om[:, None] * h / (1 + t + (sigma + tau[:, None]) * h), 0)
        sc[~live] = -np.inf
        short = np.argsort(-sc, kind='stable')[:B]
        g = []
        for c in short:
            b = np.zeros(n)
            b[cand[c, 0]], b[cand[c, 1]] = 1, -1
            g.append(sum(w * (A @ b) @ (A @ b) / (1 + b @ A @ b) for w, A in zip(om, inv)))
        j = int(np.argmax(g))
        ech = int(short[j])
        seq.append(ech)
        marg.append(g[j])
        live[ech] = False
        a, bb = cand[ech]
        Fc[a] += Gc[ech]
        Fc[bb] -= Gc[ech]
        M[a, a] += 1
        M[bb, bb] += 1
        M[a, bb] -= 1
        M[bb, a] -= 1
    assert out['sequence'] == seq
    assert np.allclose(out['marginals'], marg, rtol=1e-8)
    assert np.allclose(out['trace_estimate'], tr, rtol=1e-8)


def test_true_marginals_match_eigenvalues():
    """validate.py:26-43: certified marginals (eps=1e-9 rule, tol 1e-10) vs eigenvalue prefix gains,
    plus the twin closed form x^-s - (x+2)^-s (exact.py:20-28) on a graph with twin vertices."""
    n = 200
    e = graphs.barabasi_albert(n, 3, 11)
    G = O.Graph(n, e, 0.5)
    cand = graphs.candidates(n, e, 100, 4)
    seq = O.random_set(cand, 4, 2)
    g, cert = O.true_marginals(G, cand, seq, 0.5)
    exact = np.diff(np.r_[0.0, O.eval_prefix(G.M0.toarray(), cand, seq, 0.5)])
    assert np.max(np.abs(g - exact) / exact) < 1e-7
    # twins: two leaves hanging off the same vertex
    e2 = np.array([[0, 1], [1, 2], [2, 3], [3, 0], [0, 4], [0, 5]])
    G2 = O.Graph(6, e2, 0.5)
    cand2 = np.array([[4, 5]])
    g2, _ = O.true_marginals(G2, cand2, [0], 0.5)
    x = 1 + 0.5
    assert abs(g2[0] - (x ** -0.5 - (x + 2) ** -0.5)) < 1e-8


def test_rademacher_sketch_properties():
    n = 300
    e = graphs.barabasi_albert(n, 2, 1)
    G = O.Graph(n, e, 0.1)
    cand = graphs.candidates(n, e, 50, 1)
    Z, F, Gc = O.sketches(G, cand, 16, 9, kind='rademacher')
    assert set(np.unique(Z * 4)) == {-1.0, 1.0} and set(np.unique(Gc * 4)) == {-1.0, 1.0}
    # F = C0^T U: row sums of F over all vertices vanish (each edge contributes +U and -U)
    assert np.allclose(F.sum(0), 0, atol=1e-12)
    assert abs(np.mean(Z)) < 0.02
    Z2, _, _ = O.sketches(G, cand, 16, 10, kind='rademacher')
    assert not np.array_equal(Z, Z2)


# ---------------------------------------------------------------- pinned arithmetic (arith='cr')
def test_ddsum_is_correctly_rounded():
    import math
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 1000, 4097):
        P = rng.standard_normal((n, 3)) * np.exp(rng.uniform(-20, 20, (n, 3)))
        got = O.ddsum(P)
        for c in range(3):
            assert got[c] == math.fsum(P[:, c])


def test_cr_operator_matches_dense_and_numpy_mode():
    n, e, w = _problem()
    hub = np.array([[0, j] for j in range(1, 299)])          # one row longer than 256
    e = graphs.simple_edges(np.concatenate([e, hub]), n)
    G = O.Graph(n, e, 0.2)
    opn, opc = O.Op(G), O.Op(G, arith='cr')
    for op in (opn, opc):
        op.add(3, 200)
        op.add(200, 9)
    X = np.random.default_rng(1).standard_normal((n, 4))
    assert np.allclose(opc(X, 0.3), opn(X, 0.3), rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize('idx', [0, 1])
def test_cr_mode_reproduces_reference_fixture_sequences(idx):
    """The pinned arithmetic selects the fixture's sequences too (spmv within 1%: a different
    rounding of a chaotic recurrence may move a column's stop by an iteration)."""
    fx = FIX['configurations'][idx]
    d = SI.make(fx['n'], fx['topology'], fx['support'])
    tau, om = SI.quad(0.5, 20)
    G = O.Graph(d['n'], d['edges'], 1.0)
    for naive, name in ((True, 'Naive-ShiftedCG'), (False, 'Inc-Full')):
        out = O.polar(G, d['cand'], 15, 0.5, B=1, tol=1e-7, sketch=(d['Z'], d['F'], d['G']), naive=naive,
                      poles=(tau, om), width=48, freeze=False, arith='cr')
        ref = fx['methods'][name]
        assert out['sequence'] == ref['sequence'], name
        spmv = out['spmv'] if naive else sum(out['solves'][:-1])
        assert abs(spmv - ref['spmvs']) <= 0.01 * ref['spmvs']


def test_cr_mode_matches_dense_closed_form():
    n = 250
    e = graphs.barabasi_albert(n, 3, 7)
    G = O.Graph(n, e, 0.1)
    cand = graphs.candidates(n, e, 400, 8)
    sk = O.sketches(G, cand, 10, 3, kind='rademacher')
    tau, om, _ = O.fixed_rule(0.5, 10, 0.1, G.Lam(5))
    a = O.polar(G, cand, 5, 0.5, q=10, B=3, tol=1e-12, sketch=sk, poles=(tau, om))
    b = O.polar(G, cand, 5, 0.5, q=10, B=3, tol=1e-12, sketch=sk, poles=(tau, om), arith='cr')
    assert a['sequence'] == b['sequence']
    assert np.allclose(a['marginals'], b['marginals'], rtol=1e-9)
    assert np.allclose(a['trace_estimate'], b['trace_estimate'], rtol=1e-9)


def test_lanczos_quadrature_exact_with_full_krylov_space():
    """m-step Lanczos quadrature is the Gauss rule of the spectral measure of z: with m >= n steps
    it is exact, z^T A^-s z from the eigendecomposition (pins the oracle's slq.py restatement)."""
    n = 14
    rng = np.random.default_rng(3)
    e = graphs.barabasi_albert(n, 2, 1)
    G = O.Graph(n, e, 0.3)
    A = G.M0.toarray()
    lam, V = np.linalg.eigh(A)
    Z = rng.choice([-1.0, 1.0], size=(n, 5))
    f = lambda x: x ** -0.5
    exact = np.einsum('ic,ij,j,kj,kc->c', Z, V, f(lam), V, Z)
    got = O.lanczos_quad(lambda X: A @ X, Z, n, f)
    assert np.allclose(got, exact, rtol=1e-9, atol=0)
    # fewer steps: the Gauss rule of degree 2m-1 converges from below for this completely
    # monotone f (a dropped term or a wrong sign would not), not exact
    err = [np.max(np.abs(O.lanczos_quad(lambda X: A @ X, Z, m, f) / exact - 1)) for m in (6, 10)]
    assert 1e-5 < err[0] < 1e-2 and err[1] < 1e-5
    assert np.all(O.lanczos_quad(lambda X: A @ X, Z, 6, f) <= exact)


def test_slq_greedy_estimates_are_trace_differences():
    """With m = n the SLQ estimate of candidate e is exactly mean_z z^T (f(M) - f(M + b b^T)) z."""
    n = 16
    e = graphs.barabasi_albert(n, 2, 4)
    cand = graphs.candidates(n, e, 6, 2)
    G = O.Graph(n, e, 0.2)
    out = O.slq_greedy(G, cand, 2, 0.5, probes=4, m=n, chunk=8, seed=5)
    A = G.M0.toarray()
    Z = np.random.default_rng(5).choice([-1.0, 1.0], size=(n, 4))
    def fq(M):
        lam, V = np.linalg.eigh(M)
        return np.mean(np.einsum('ic,ij,j,kj,kc->c', Z, V, lam ** -0.5, V, Z))
    base = fq(A)
    want = []
    for a, b in cand:
        Mb = A.copy(); Mb[a, a] += 1; Mb[b, b] += 1; Mb[a, b] -= 1; Mb[b, a] -= 1
        want.append(base - fq(Mb))
    assert np.allclose(out['estimates'][0], want, rtol=1e-8, atol=1e-12)
    assert out['sequence'][0] == int(np.argmax(want))


@pytest.mark.parametrize('obj', ['er', 'logdet'])
def test_er_logdet_marginals_match_eigenvalues(obj):
    """The single-pole objectives (polar_gpu.py:39-43: one pole at 0, weight 1): certified marginals
    and the B > 1 greedy's re-rank marginals equal the dense prefix gains, sum 1/lambda (er) and
    log det (logdet, the matrix determinant lemma log(1 + b^T M^-1 b))."""
    n = 150
    e = graphs.barabasi_albert(n, 3, 21)
    G = O.Graph(n, e, 0.3)
    cand = graphs.candidates(n, e, 80, 6)
    seq = O.random_set(cand, 4, 3)
    g, cert = O.true_marginals(G, cand, seq, 0.5, obj=obj)
    exact = np.diff(np.r_[0.0, O.eval_prefix(G.M0.toarray(), cand, seq, 0.5, obj=obj)])
    assert cert['p'] == 1
    assert np.max(np.abs(g - exact) / np.abs(exact)) < 1e-8
    r = O.polar(G, cand, 4, 0.5, q=16, B=4, tol=1e-10, obj=obj, kind='rademacher', seed=5)
    ex = np.diff(np.r_[0.0, O.eval_prefix(G.M0.toarray(), cand, r['sequence'], 0.5, obj=obj)])
    assert np.max(np.abs(np.array(r['marginals']) - ex) / np.abs(ex)) < 1e-8


def test_polar_column_parallel_equals_serial():
    """The CPU baseline splits every block solve's columns over forked processes; columns are
    independent, so sequence and counters equal the one-process run; with the pinned arithmetic
    (order-independent reductions) the marginals and trace estimates are bit-identical, with NumPy's
    native reductions (whose order follows the block shape) within 1e-12."""
    c = configs.make('c2', 0.01)
    Go = O.Graph(c['n'], c['edges'], c['sigma'])
    sk = O.sketches(Go, c['cand'], 16, c['seed'], kind='rademacher')
    kw = dict(q=16, B=4, tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']), width=32)
    for arith, bound in (('cr', 0.0), ('numpy', 1e-12)):
        a = O.polar(Go, c['cand'], 4, c['s'], arith=arith, **kw)
        b = O.polar(Go, c['cand'], 4, c['s'], arith=arith, workers=3, **kw)
        assert a['sequence'] == b['sequence'] and (a['spmv'], a['rhs']) == (b['spmv'], b['rhs'])
        assert a['solves'] == b['solves']
        for key in ('marginals', 'trace_estimate'):
            x, y = np.asarray(a[key]), np.asarray(b[key])
            assert np.max(np.abs(x - y) / np.abs(x)) <= bound, (arith, key)
