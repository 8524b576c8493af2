"""The adjoint, batched and full-scale configurations at their workload sizes.

* c5 (Chung-Lu n=4M, m=34.7M, sigma=0.02, s=0.5, 32 poles, N=1e7): the real graph and pool with a
  bounded greedy -- q=8 (a bank of 16 columns), k=3, B=16 -- in FP64 and FP32-mixed.  Per-round
  marginals agree within 1e-5 relative.  A pick joining two vertices of degree 0 (in the graph of
  that round) has the re-rank score sum_j om_j [1/(sigma+tau_j) - 1/(sigma+tau_j+2)] exactly: the
  quadrature rule's value of the closed-form gain sigma^-s - (sigma+2)^-s (u_j = b/(sigma+tau_j)
  on the two endpoints); checked within 1e-6, and the rule's value against the closed form within
  the rule's own error at sigma and sigma+2.
* c3 (weighted Chung-Lu n=1M, 2M candidates, FP32-mixed): forward + backward of the sketched
  trace over all 10M pairs, checked on 200 sampled pairs by central differences (h=1e-4) through
  Sherman-Morrison from independent FP64 solves (polar_amd.run_adjoint).
* c4 (64 R-MAT graphs, n log-uniform in [1e4, 5e5], FP32-mixed, k=10, B=4): the batch in one
  launch sequence against polar_gpu.polar per graph, bit for bit (selections, marginals, trace
  estimates, counters)."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from polar_amd import configs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _say(*a):
    print('[fullsize]', *a, file=sys.stderr, flush=True)


def test_c5_bounded_greedy_fp32_vs_fp64_and_isolated_pairs():
    _say('c5: building graph and pool')
    c = configs.make('c5')
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    tau, om = c['tau'], c['om']
    sk = pg.sketches(G, c['cand'], 8, c['seed'], kind='rademacher', on_device='lazy')
    kw = dict(q=8, B=16, tol=c['tol'], sketch=sk, poles=(tau, om))
    runs = {}
    for prec in ('fp64', 'fp32'):
        _say(f'c5: k=3 greedy {prec}')
        runs[prec] = pg.polar(G, c['cand'], 3, c['s'], precision=prec, **kw)
        torch.cuda.empty_cache()
    r64, r32 = runs['fp64'], runs['fp32']
    assert r32['replace'] is True and len(r64['marginals']) == len(r32['marginals']) == 3
    m64, m32 = np.array(r64['marginals']), np.array(r32['marginals'])
    assert np.max(np.abs(m32 - m64) / np.abs(m64)) <= 1e-5
    sig, s = c['sigma'], c['s']
    rule = float(np.sum(om * (1.0 / (sig + tau) - 1.0 / (sig + tau + 2.0))))
    closed = sig ** -s - (sig + 2.0) ** -s
    err = sum(abs(float(np.sum(om / (x + tau))) - x ** -s) for x in (sig, sig + 2.0))
    assert abs(rule - closed) <= err * (1 + 1e-12)
    deg = np.bincount(c['edges'][:, 0], minlength=c['n']) + np.bincount(c['edges'][:, 1], minlength=c['n'])
    checked = 0
    for r in (r64, r32):
        used = set()
        for e, g in zip(r['sequence'], r['marginals']):
            a, b = (int(x) for x in c['cand'][e])
            if deg[a] == 0 and deg[b] == 0 and a not in used and b not in used:
                assert abs(g - rule) <= 1e-6 * rule, (e, g, rule)
                checked += 1
            used.update((a, b))
    _say(f'c5: isolated-pair picks checked: {checked}; rule {rule:.8g}, closed form {closed:.8g}')


def test_c3_adjoint_full_scale_central_differences():
    from polar_amd import run_adjoint
    _say('c3: forward + backward over all pairs, 200 central differences')
    res = run_adjoint.main(['--skip-greedy', '--samples', '200'])
    assert res['pairs'] == res['workload']['m'] + res['workload']['N'] == 10_000_000
    chk = res['check']
    assert chk['samples'] == 200 and chk['max_rel_err'] <= 1e-5, chk
    torch.cuda.empty_cache()


def test_c4_batch_of_64_rmat_graphs_equals_per_graph():
    from polar_amd import batched, run_batched
    _say('c4: building 64 R-MAT graphs')
    inp = run_batched.build_inputs(64, 1.0)
    Gs = [pg.Graph(x['n'], x['edges'], x['sigma']) for x in inp]
    cands = [x['cand'] for x in inp]
    poles = [(x['tau'], x['om']) for x in inp]
    s_all = [x['s'] for x in inp]
    k, B, q = inp[0]['k'], inp[0]['B'], inp[0]['q']

    def sks():
        return [pg.sketches(G, cands[g], q, inp[g]['seed'], kind='rademacher', on_device='lazy')
                for g, G in enumerate(Gs)]
    _say('c4: batched')
    out = batched.polar_batched(Gs, cands, k, s_all, q=q, B=B, sketch=sks(), poles=poles, precision='fp32',
                                graph=False)
    sk = sks()
    for g, G in enumerate(Gs):
        ref = pg.polar(G, cands[g], k, s_all[g], q=q, B=B, sketch=sk[g], poles=poles[g], precision='fp32',
                       history=False)
        got = out[g]
        assert got['sequence'] == ref['sequence'], g
        assert got['marginals'] == ref['marginals'], g
        assert got['trace_estimate'] == ref['trace_estimate'], g
        assert (got['spmv'], got['rhs']) == (ref['spmv'], ref['rhs']), g
    _say(f'c4: 64 graphs equal; n_total={sum(G.n for G in Gs)}')
