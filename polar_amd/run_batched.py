"""Batched-form CLI: the 64-graph R-MAT workload run as one batch, against the one-graph-at-a-time
loop of the reference's harness (realworld/run_real.py:114-133 calls polar_gpu.polar per
graph).

    python -m polar_amd.run_batched [--graphs 64] [--scale 1] [--precision fp32] [--tune]
                                    [--modes batched nograph sequential] [--out results/batched_c4.json]

Workload (configs.rmat_spec): graph i has n log-uniform in [1e4, 5e5] (seed 100+i), R-MAT
(0.57, 0.19, 0.19, 0.05), edge factor 8, unit weights; sigma=0.05, s drawn from {.25,.5,.75},
q=32 Rademacher (seed 100+i), 16 poles (fixed_rule), N=2,000 candidates, k=10, B=4, FP32-mixed.
Modes: `batched` (polar_batched, hipGraph replay), `nograph` (polar_batched, direct launches),
`sequential` (polar_gpu.polar per graph; its seconds are the sum of the per-graph timed regions).
Every mode must select the same edges; the JSON reports times, column-iterations and the check.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time
from pathlib import Path

import numpy as np


def _make(args):
    from . import configs
    i, scale = args
    return configs.make(f'c4_{i}', scale)


def build_inputs(count, scale, workers=None):
    """Host inputs of graphs 0..count-1, generated in worker processes (forked before any GPU use;
    started as fresh interpreters once this process has initialised the GPU runtime)."""
    workers = workers or min(16, os.cpu_count() or 1, count)
    torch = sys.modules.get('torch')
    ctx = 'spawn' if torch is not None and torch.cuda.is_initialized() else 'fork'
    jobs = [(i, scale) for i in range(count)]
    if workers <= 1:
        return [_make(j) for j in jobs]
    with mp.get_context(ctx).Pool(workers) as pool:
        return pool.map(_make, jobs, chunksize=1)


def log(*a):
    print('[batched]', *a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--graphs', type=int, default=64)
    ap.add_argument('--scale', type=float, default=1.0)
    ap.add_argument('--precision', choices=['fp32', 'fp64'], default='fp32')
    ap.add_argument('--modes', nargs='+', default=['batched', 'nograph', 'sequential'])
    ap.add_argument('--tune', action='store_true', help='autotune the batched launch shape first')
    ap.add_argument('--k', type=int, default=None)
    ap.add_argument('--out', default=None)
    a = ap.parse_args(argv)
    t = time.time()
    inp = build_inputs(a.graphs, a.scale)
    log(f'{a.graphs} graphs, n={sum(c["n"] for c in inp)} m={sum(c["m"] for c in inp)} built in {time.time() - t:.1f}s')

    import torch
    from . import autotune, batched
    from . import polar_gpu as pg
    Gs = [pg.Graph(c['n'], c['edges'], c['sigma'], weights=c['weights']) for c in inp]
    cands = [c['cand'] for c in inp]
    poles = [(c['tau'], c['om']) for c in inp]
    s_all = [c['s'] for c in inp]
    k = a.k or inp[0]['k']
    B, q = inp[0]['B'], inp[0]['q']

    def sks():
        return [pg.sketches(G, cands[g], q, inp[g]['seed'], kind='rademacher', on_device='lazy')
                for g, G in enumerate(Gs)]

    tune = None
    res = dict(workload=dict(name='c4', graphs=a.graphs, scale=a.scale, n_total=int(sum(c['n'] for c in inp)),
                             m_total=int(sum(c['m'] for c in inp)), k=k, B=B, q=q, p=len(poles[0][0]),
                             sigma=inp[0]['sigma'], s=s_all, precision=a.precision),
               device=torch.cuda.get_device_name(0), modes={})
    if a.tune:
        U = batched.UnionGraph(Gs)
        op = pg.Op(U, precision=a.precision)
        actives = [np.unique(c) for c in cands]
        a0 = np.r_[0, np.cumsum([len(x) for x in actives])]
        segs = batched.Segments(*U.segments(a0))
        act = torch.from_numpy(np.concatenate([x + U.row0[g] for g, x in enumerate(actives)]).astype(np.int32)).cuda()
        t = time.time()
        tune = autotune.tune(op, 2 * q, np.stack([p[0] for p in poles]), act, segs=segs, log=log)
        res['tune'] = dict(tune, seconds=time.time() - t)
        log('tuned', tune)
        del op, U
    seqs = {}
    warm = False
    for mode in a.modes:
        if mode in ('batched', 'nograph'):
            kw = dict(graph=mode == 'batched')
            if tune:
                kw['tune'] = dict(tune, graph=mode == 'batched')
            if not warm:                       # first use of the kernels and the allocator
                batched.polar_batched(Gs, cands, k, s_all, q=q, B=B, sketch=sks(), poles=poles,
                                      precision=a.precision, **kw)
                warm = True
            t = time.time()
            out = batched.polar_batched(Gs, cands, k, s_all, q=q, B=B, sketch=sks(), poles=poles,
                                        precision=a.precision, **kw)
            wall = time.time() - t
            seconds = out[0]['seconds']
        elif mode == 'sequential':
            out = []
            sk = sks()
            t = time.time()
            for g, G in enumerate(Gs):
                out.append(pg.polar(G, cands[g], k, s_all[g], q=q, B=B, sketch=sk[g], poles=poles[g],
                                    precision=a.precision))
            wall = time.time() - t
            seconds = float(sum(o['seconds'] for o in out))
        else:
            raise ValueError(mode)
        seqs[mode] = [o['sequence'] for o in out]
        res['modes'][mode] = dict(seconds=seconds, wall_seconds=wall,
                                  column_iterations=int(sum(o['spmv'] for o in out)),
                                  rhs=int(sum(o['rhs'] for o in out)))
        log(mode, f'{seconds:.3f}s', res['modes'][mode])
    first = next(iter(seqs.values()))
    res['selections_identical'] = all(v == first for v in seqs.values())
    if 'sequential' in res['modes'] and 'batched' in res['modes']:
        res['speedup_batched_vs_sequential'] = res['modes']['sequential']['seconds'] / res['modes']['batched']['seconds']
    if 'nograph' in res['modes'] and 'batched' in res['modes']:
        res['speedup_graph_vs_direct'] = res['modes']['nograph']['seconds'] / res['modes']['batched']['seconds']
    print(json.dumps(res))
    if a.out:
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps(res, indent=1))
    return res


if __name__ == '__main__':
    main()
