"""Adjoint form: dT/dw from the forward's shifted blocks vs dense algebra, central differences
(h=1e-4) and the bank's h cache (T = sketched fractional trace, see polar_amd/adjoint.py)."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import adjoint, graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _problem():
    n = 300
    e = graphs.barabasi_albert(n, 3, 2)
    rng = np.random.default_rng(0)
    w = rng.uniform(0.5, 1.5, len(e))
    cand = graphs.candidates(n, e, 40, 5)
    pairs = np.concatenate([e, cand])
    wall = np.concatenate([w, np.zeros(len(cand))])
    sigma = 0.2
    G0 = O.Graph(n, e, sigma, weights=w)
    tau, om, _ = pg.fixed_rule(0.5, 10, sigma, G0.Lam(0))
    Z = O.rademacher_signs(4, 0, np.arange(n), 12) / np.sqrt(12)
    return n, e, w, cand, pairs, wall, sigma, tau, om, Z, G0


def test_gradient_matches_dense_and_autograd():
    n, e, w, cand, pairs, wall, sigma, tau, om, Z, G0 = _problem()
    M = G0.M0.toarray()
    inv = [np.linalg.inv(M + t * np.eye(n)) for t in tau]
    T_dense = sum(o * np.trace(Z.T @ A @ Z) for o, A in zip(om, inv))
    g_dense = np.array([-sum(o * np.sum(((A[a] - A[b]) @ Z) ** 2) for o, A in zip(om, inv)) for a, b in pairs])
    wt = torch.tensor(wall, dtype=torch.float64, requires_grad=True)
    T = adjoint.SketchedTrace.apply(wt, n, pairs, sigma, tau, om, Z, 1e-12)
    T.backward()
    assert abs(T.item() - T_dense) <= 1e-9 * abs(T_dense)
    assert np.max(np.abs(wt.grad.numpy() - g_dense)) <= 1e-8 * np.max(np.abs(g_dense))
    assert np.all(g_dense < 0)                       # adding weight always lowers the trace


def test_gradient_central_differences_and_proxy():
    n, e, w, cand, pairs, wall, sigma, tau, om, Z, G0 = _problem()
    T0, g = adjoint.sketched_trace(n, pairs, wall, sigma, tau, om, Z, tol=1e-12)
    g = g.cpu().numpy()
    rng = np.random.default_rng(3)
    h = 1e-4
    for i in rng.choice(len(pairs), 12, replace=False):
        wp, wm = wall.copy(), wall.copy()
        wp[i] += h
        wm[i] -= h if wall[i] > h else 0.0
        Tp, _ = adjoint.sketched_trace(n, pairs, wp, sigma, tau, om, Z, tol=1e-12, want_grad=False)
        Tm, _ = adjoint.sketched_trace(n, pairs, wm, sigma, tau, om, Z, tol=1e-12, want_grad=False)
        fd = (Tp - Tm) / (wp[i] - wm[i])
        assert abs(fd - g[i]) <= 1e-5 * abs(g[i]) + 1e-9, (i, fd, g[i])
    # first-order proxy == pole-weighted h cache of the bank (the reference's proxy numerator)
    Xo = O.msolve(O.Op(G0), Z, tau, np.arange(n), 1e-12)[0]
    hc = np.sum((Xo[:, cand[:, 0]] - Xo[:, cand[:, 1]]) ** 2, axis=-1)
    proxy = np.sum(om[:, None] * hc, axis=0)
    gc = -g[len(e):]
    assert np.max(np.abs(gc - proxy)) <= 1e-9 * np.max(proxy)
    sl = adjoint.gradient_shortlist(torch.from_numpy(g[len(e):]), 5)
    assert sl == np.argsort(-proxy, kind='stable')[:5].tolist()


def test_fp32_gradient_error():
    n, e, w, cand, pairs, wall, sigma, tau, om, Z, G0 = _problem()
    T64, g64 = adjoint.sketched_trace(n, pairs, wall, sigma, tau, om, Z, tol=1e-10)
    T32, g32 = adjoint.sketched_trace(n, pairs, wall, sigma, tau, om, Z, tol=1e-7, precision='fp32')
    g64, g32 = g64.cpu().numpy(), g32.cpu().numpy()
    assert abs(T32 - T64) <= 1e-5 * abs(T64)
    assert np.max(np.abs(g32 - g64)) <= 1e-4 * np.max(np.abs(g64))


def test_gradient_shortlist_is_first_round_shortlist():
    """polar(shortlist='gradient') ranks round 1 by sum_j om_j h[j,e], the first-order gain: the
    same order as -dT/dw over the candidates from the adjoint."""
    n = 600
    e = graphs.barabasi_albert(n, 3, 4)
    cand = graphs.candidates(n, e, 80, 6)
    sigma = 0.1
    G = pg.Graph(n, e, sigma)
    tau, om, _ = pg.fixed_rule(0.5, 10, sigma, G.Lam(3))
    q = 16
    sk = pg.sketches(G, cand, q, 7, kind='rademacher', on_device=True)
    tr = []
    out = pg.polar(G, cand, 3, 0.5, q=q, B=6, sketch=sk, poles=(tau, om), tol=1e-12, shortlist='gradient', trace=tr)
    pairs = np.concatenate([e, cand])
    wall = np.concatenate([np.ones(len(e)), np.zeros(len(cand))])
    _, g = adjoint.sketched_trace(n, pairs, wall, sigma, tau, om, sk[0], tol=1e-12)
    assert tr[0]['shortlist'] == adjoint.gradient_shortlist(g[len(e):], 6)
    assert len(out['sequence']) == 3


def test_run_adjoint_cli_small():
    from polar_amd import run_adjoint
    res = run_adjoint.main(['--scale', '0.004', '--samples', '24', '--k', '3', '--B', '4', '--precision', 'fp64',
                            '--tol', '1e-11'])
    assert res['check']['max_rel_err'] <= 1e-7
    res32 = run_adjoint.main(['--scale', '0.004', '--samples', '24', '--skip-greedy'])
    assert res32['check']['max_rel_err'] <= 1e-3
    assert set(res['greedy']['runs']) == {'gradient', 'proxy'}
