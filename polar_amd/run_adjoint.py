"""Adjoint-form CLI on the weighted Chung-Lu workload (configs 'c3').

    python -m polar_amd.run_adjoint [--scale 1] [--samples 200] [--h 1e-4] [--precision fp32]
                                    [--k 20] [--B 16] [--out results/adjoint_c3.json]

Workload: n=1,000,000, gamma 2.3, mean degree 16 (~8M edges), weights U[0.5, 1.5], seed 2;
sigma=0.05, s=0.5, q=64 Rademacher columns, 24 poles; the 2,000,000 candidate non-edges enter at
weight 0.

1. Forward + backward: T = SketchedTrace(w) over all edges and candidates, T.backward() gives
   dT/dw for every pair (edges and candidates) from the forward's shifted blocks.
2. Check on `samples` pairs (half edges, half candidates): central differences with step h,
       CD = [T(w + h e) - T(w - h e)] / 2h = -sum_j om_j g_j / (1 - h^2 r_j^2),
   g_j = ||Z^T u_j||^2, r_j = b^T u_j, u_j = (M + tau_j)^{-1} b, evaluated exactly through the
   Sherman-Morrison identity from independent FP64 solves of the pair vectors b = e_a - e_b (tol
   1e-12): no perturbed re-solve of the whole system, whose CG tolerance would swamp a 1e-4 step.
3. Greedy: k rounds with the gradient-ranked top-B shortlist (first-order gain) and, beside it,
   with the reference's proxy shortlist; both re-rank by exact marginals.
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def log(*a):
    print('[adjoint]', *a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--scale', type=float, default=1.0)
    ap.add_argument('--samples', type=int, default=200)
    ap.add_argument('--h', type=float, default=1e-4)
    ap.add_argument('--precision', choices=['fp32', 'fp64'], default='fp32')
    ap.add_argument('--tol', type=float, default=1e-7)
    ap.add_argument('--k', type=int, default=None)
    ap.add_argument('--B', type=int, default=None)
    ap.add_argument('--skip-greedy', action='store_true')
    ap.add_argument('--out', default=None)
    a = ap.parse_args(argv)

    from . import configs
    t = time.time()
    c = configs.make('c3', a.scale)
    n, E, w, cand = c['n'], c['edges'], c['weights'], c['cand']
    sigma, tau, om, q, seed = c['sigma'], c['tau'], c['om'], c['q'], c['seed']
    k = a.k or c['k']
    B = a.B or c['B']
    log(f'c3 n={n} m={len(E)} N={len(cand)} built in {time.time() - t:.1f}s')

    import torch
    from . import adjoint
    from . import polar_gpu as pg
    G = pg.Graph(n, E, sigma, weights=w)
    Z = pg.sketches(G, cand, q, seed, kind='rademacher', on_device=True, precision='fp64')[0]
    pairs = np.concatenate([E, cand])
    wall = np.concatenate([w, np.zeros(len(cand))])
    res = dict(workload=dict(name='c3', n=int(n), m=int(len(E)), N=int(len(cand)), sigma=sigma, s=c['s'], q=q,
                             p=len(tau), precision=a.precision, tol=a.tol, scale=a.scale),
               device=torch.cuda.get_device_name(0))

    # 1. forward + backward
    wt = torch.tensor(wall, dtype=torch.float64, device='cuda', requires_grad=True)
    torch.cuda.synchronize()
    t = time.time()
    T = adjoint.SketchedTrace.apply(wt, n, pairs, sigma, tau, om, Z, a.tol, a.precision)
    T.backward()
    torch.cuda.synchronize()
    res['forward_backward_seconds'] = time.time() - t
    grad = wt.grad.detach()
    res['T'] = float(T.item())
    res['pairs'] = int(len(pairs))
    log(f'T={res["T"]:.10g} forward+backward over {len(pairs)} pairs in {res["forward_backward_seconds"]:.2f}s')

    # 2. central differences through Sherman-Morrison on sampled edges and candidates
    rng = np.random.default_rng(2026)
    ns = a.samples
    idx = np.concatenate([rng.choice(len(E), ns // 2, replace=False),
                          len(E) + rng.choice(len(cand), ns - ns // 2, replace=False)])
    ab = pairs[idx]
    Wb = torch.zeros((n, ns), dtype=torch.float64, device='cuda')
    cols = torch.arange(ns, device='cuda')
    Wb[torch.from_numpy(ab[:, 0]).cuda(), cols] = 1.0
    Wb[torch.from_numpy(ab[:, 1]).cuda(), cols] = -1.0
    t = time.time()
    Uj = pg.msolve(pg.Op(G), Wb, tau, np.arange(n), 1e-12, block=True)[0]          # (p, n, ns)
    Zd = Z.to(torch.float64)
    cd = torch.zeros(ns, dtype=torch.float64, device='cuda')
    ia = torch.from_numpy(ab[:, 0]).cuda()
    ib = torch.from_numpy(ab[:, 1]).cuda()
    for j in range(len(tau)):
        g = (Zd.T @ Uj[j]).pow(2).sum(0)
        r = Uj[j, ia, cols] - Uj[j, ib, cols]
        cd -= om[j] * g / (1.0 - (a.h * r) ** 2)
    torch.cuda.synchronize()
    del Uj
    got = grad[torch.from_numpy(idx).cuda()]
    rel = ((got - cd).abs() / cd.abs()).cpu().numpy()
    res['check'] = dict(samples=ns, h=a.h, max_rel_err=float(rel.max()), median_rel_err=float(np.median(rel)),
                        max_rel_err_edges=float(rel[:ns // 2].max()), max_rel_err_candidates=float(rel[ns // 2:].max()),
                        seconds=time.time() - t,
                        method='CD through Sherman-Morrison from FP64 solves of e_a - e_b (tol 1e-12)')
    log('check', res['check'])

    # 3. gradient-ranked shortlist fed to the greedy rounds, beside the reference's proxy
    if not a.skip_greedy:
        sk = pg.sketches(G, cand, q, seed, kind='rademacher', on_device='lazy')
        runs = {}
        for mode in ('gradient', 'proxy'):
            out = pg.polar(G, cand, k, c['s'], q=q, B=B, tol=a.tol, sketch=sk, poles=(tau, om),
                           precision=a.precision, shortlist=mode)
            runs[mode] = dict(seconds=out['seconds'], sequence=out['sequence'], marginals=out['marginals'],
                              gain=float(np.sum(out['marginals'])), spmv=out['spmv'], rhs=out['rhs'])
            log(mode, f'{out["seconds"]:.2f}s gain={runs[mode]["gain"]:.6g}')
        res['greedy'] = dict(k=k, B=B, runs=runs,
                             common_selections=len(set(runs['gradient']['sequence']) & set(runs['proxy']['sequence'])))
    print(json.dumps({kk: v for kk, v in res.items() if kk != 'greedy'} |
                     ({'greedy_gain': {m: r['gain'] for m, r in res['greedy']['runs'].items()}} if 'greedy' in res else {})))
    if a.out:
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps(res, indent=1))
    return res


if __name__ == '__main__':
    main()
