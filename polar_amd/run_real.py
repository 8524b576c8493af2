"""Drop-in for realworld/run_real.py (`time` and `effect` modes) on the MI355X core.

    python -m polar_amd.run_real time   --graphs c2 --s 0.5 --poles 24 --sigma 0.05
    python -m polar_amd.run_real effect --graphs c1 --s 0.5 --poles 16 --sigma 0.1 --k 10 --B 1

Same CLI flags, JSON schema and file names as the reference (run_real.py:23-133, 174-192):
`time_<graph>.json` = {mode, graph, params, env, runs{name: [polar dict]}};
`effect_<graph>.json` = {mode, graph, params, env, selections, gains{target: {label: [...],
'_evaluator': {...}}}, seconds_total}.  Additions: --sigma (the reference's module global
SIGMA), --poles P (fixed-count rule instead of `composite`), --kind (sketch kind), --precision,
--out (results directory) and synthetic graph names (c1, c2, c3, c5, c4_<i>, see configs.py) in
--graphs; SNAP/KONECT names load from $POLAR_DATA like the reference.  --slq adds the SLQ-Greedy
baseline (polar_amd/slq.py) as in the reference.  `opt` and `param` modes are not part of this
build.
"""
import argparse
import json
import platform
import time
import zlib
from pathlib import Path

import numpy as np
import torch

from . import configs, exact, graphs
from . import polar_gpu as pg

OUT = Path.cwd() / 'results'
SUF = ''
SIGMA = 1.0
POOL = 'uniform'
POOL_GEN = 'reference'
PSEED = 0


def env():
    """run_real.py:23-25."""
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else 'none'
    return dict(gpu=name, torch=torch.__version__, numpy=np.__version__, python=platform.python_version(),
                host=platform.node())


def setup(name, N, seed_offset=None, sigma=None):
    """run_real.py:32-45; synthetic config names build their graph and pool with configs.make.

    N=None: the reference's default of 2000 candidates for a named graph, the configuration's own
    pool for a synthetic one; N=0: every non-edge (named graph) or the configuration's pool; N > 0:
    that many (a synthetic pool keeps its first N).  Named graphs draw the uniform pool with the
    reference's generator (graphs.candidates, realworld/graphs.py:61-69) unless POOL_GEN is
    'fast' (graphs.candidates_fast: same distribution, vectorised, its own draw order); meta
    records which."""
    sigma = SIGMA if sigma is None else sigma
    seed_offset = PSEED if seed_offset is None else seed_offset
    if name in configs.SPECS or name.startswith('c4_'):
        c = configs.make(name)
        n, E = c['n'], c['edges']
        G = pg.Graph(n, E, sigma, weights=c['weights'])
        cand = c['cand'] if not N else c['cand'][:N]
        meta = dict(name=name, source='synthetic', n=int(n), m=int(len(E)),
                    preprocessing='generated (configs.py): undirected, deduplicated, no self-loops',
                    pool_generator='configs.make')
        seed = c['seed']
    else:
        n, E, meta = graphs.load(name)
        G = pg.Graph(n, E, sigma)
        seed = 20270000 + zlib.crc32(name.encode()) % 10000 + seed_offset
        N = 2000 if N is None else N
        if not N:
            cand, gen = graphs.all_nonedges(n, E), 'all_nonedges'
        elif POOL == '2hop':
            cand, gen = graphs.candidates_2hop(n, E, N, seed), 'candidates_2hop'
        elif POOL_GEN == 'fast':
            cand, gen = graphs.candidates_fast(n, E, N, seed), 'candidates_fast'
        else:
            cand, gen = graphs.candidates(n, E, N, seed), 'candidates'
        meta['pool_generator'] = gen
    meta.update(N=len(cand), candidate_seed=seed, pool=POOL, dmax=G.dmax, sigma=sigma,
                n_active=int(len(np.unique(cand))))
    return G, cand, meta


def dense(G):
    """run_real.py:48-49."""
    return G.M0.get().toarray()


def _poles(a, G, k, s, obj):
    if obj != 'frac' or not a.poles:
        return None
    tau, om, _ = pg.fixed_rule(s, a.poles, G.sigma, G.Lam(k))
    return tau, om


def evaluate(G, cand, seq, s, obj, use_dense, Md):
    """run_real.py:52-56."""
    if use_dense:
        return exact.eval_prefix(Md, cand, seq, s, obj).tolist(), dict(evaluator='eigenvalues')
    g, cert = pg.true_marginals(G, cand, seq, s, obj=obj)
    return np.cumsum(g).tolist(), dict(evaluator='certified-quadrature', **cert)


def _sk(a, G, cand, seed=2026):
    if a.kind == 'rademacher':
        return pg.sketches(G, cand, a.q, seed, kind='rademacher', on_device='lazy')
    return pg.sketches(G, cand, a.q, seed, kind=a.kind)


def _progress(a, name, label):
    if not a.progress:
        return None

    def show(stage, info):
        if stage == 'bank':
            print(f"{name} {label} bank {info['seconds']:.2f}s spmv={info['spmv']}", flush=True)
        else:
            print(f"{name} {label} round {info['round']} edge={info['edge']} marginal={info['marginal']:.6g} "
                  f"{info['seconds']:.2f}s spmv={info['spmv']}", flush=True)
    return show


def _block(a):
    return {'auto': None, 'on': True, 'off': False}[getattr(a, 'block', 'auto')]


def _clean(out):
    return {k: v for k, v in out.items() if k not in ('solves',)}


def effect(a):
    """run_real.py:59-111 (Exact(s) selections need the dense exact greedy, not built here)."""
    for name in a.graphs:
        G, cand, meta = setup(name, a.N, sigma=a.sigma)
        k = a.k
        use_dense = G.n <= a.dense_max
        Md = dense(G) if use_dense else None
        sk = _sk(a, G, cand)
        res = dict(mode='effect', graph=meta, params=dict(k=k, q=a.q, B=a.B, eps=a.eps, tol=a.tol, s=a.s),
                   env=env(), selections={}, gains={})
        sel = res['selections']
        t = time.perf_counter()

        def add(label, out):
            sel[label] = _clean(out)
            print(name, label, f"{out.get('seconds', 0):.2f}s", flush=True)
        for s in a.s:
            add(f'POLAR(s={s})', pg.polar(G, cand, k, s, q=a.q, B=a.B, eps=a.eps, tol=a.tol, sketch=sk,
                                          poles=_poles(a, G, k, s, 'frac'), precision=a.precision))
        add('Sketched-ER', pg.polar(G, cand, k, 1.0, q=a.q, B=a.B, tol=a.tol, obj='er', sketch=sk,
                                    precision=a.precision))
        add('Sketched-Logdet', pg.polar(G, cand, k, 0.5, q=a.q, B=a.B, tol=a.tol, obj='logdet', sketch=sk,
                                        precision=a.precision))
        if a.slq:                                                               # run_real.py:76-79
            from . import slq
            for s in a.s:
                add(f'SLQ(s={s})', slq.slq_greedy(G, cand, k, s, probes=a.slq_probes, m=a.slq_steps))
        add('Degree', dict(sequence=pg.degree_greedy(G, cand, k)))
        for r in range(a.random):
            add(f'Random#{r}', dict(sequence=pg.random_set(cand, k, 1000 + r)))
        targets = [('frac', s) for s in a.s] + [('er', 1.0), ('logdet', 0.0)]
        keys = [f'{o}' if o != 'frac' else f'frac(s={s})' for o, s in targets]
        for key in keys:
            res['gains'][key] = {}
        out_dir = Path(a.out)
        if use_dense:
            for label, out in sel.items():
                vals = exact.eval_prefix_multi(Md, cand, out['sequence'], targets)
                for key, v in zip(keys, vals):
                    res['gains'][key][label] = v.tolist()
            for key in keys:
                res['gains'][key]['_evaluator'] = dict(evaluator='eigenvalues')
        else:
            for (obj, s), key in zip(targets, keys):
                labels = [lb for lb in sel if not lb.startswith('Random')] if (obj != 'frac' and a.cross_only) else list(sel)
                for label in labels:
                    g, info = evaluate(G, cand, sel[label]['sequence'], s, obj, False, None)
                    res['gains'][key][label] = g
                res['gains'][key]['_evaluator'] = info
                print(name, 'evaluated', key, flush=True)
                out_dir.mkdir(parents=True, exist_ok=True)
                (out_dir / f'effect_{name}{SUF}.json').write_text(json.dumps(res, indent=1))
        res['seconds_total'] = time.perf_counter() - t
        out_dir.mkdir(parents=True, exist_ok=True)
        (out_dir / f'effect_{name}{SUF}.json').write_text(json.dumps(res, indent=1))


def timing(a):
    """run_real.py:114-133: Naive vs POLAR(B=1) vs POLAR(B), rotated order, naive time limit."""
    for name in a.graphs:
        G, cand, meta = setup(name, a.N, sigma=a.sigma)
        sk = _sk(a, G, cand)
        if torch.cuda.is_available():
            free, total = torch.cuda.mem_get_info()
            meta.update(device_memory_total=int(total), device_memory_free_at_start=int(free))
        res = dict(mode='time', graph=meta, params=dict(k=a.k, q=a.q, B=a.B, eps=a.eps, tol=a.tol, s=a.s[0],
                   repetitions=a.reps, precision=a.precision, poles=a.poles or 'composite', block=a.block,
                   shortlist=a.shortlist),
                   env=env(), runs={})
        cfg = {'Naive': dict(naive=True), 'POLAR(B=1)': dict(B=1), f'POLAR(B={a.B})': dict(B=a.B)}
        if a.skip_naive:
            cfg.pop('Naive')
        if a.only:
            keep = {'Naive': 'Naive', 'B1': 'POLAR(B=1)', 'B': f'POLAR(B={a.B})'}
            cfg = {k: v for k, v in cfg.items() if k in [keep[o] for o in a.only]}
        names = list(cfg)
        reps = a.reps if meta['m'] < 2e7 else 1
        res['params'].update(naive_reps=a.naive_reps, naive_limit=a.naive_limit, repetitions=reps)
        out_dir = Path(a.out)
        for rep in range(reps):
            for nm in names[rep % len(names):] + names[:rep % len(names)]:
                if nm == 'Naive' and rep >= a.naive_reps:
                    continue
                lim = a.naive_limit if nm == 'Naive' else None
                out = pg.polar(G, cand, a.k, a.s[0], q=a.q, eps=a.eps, tol=a.tol, sketch=sk, time_limit=lim,
                               poles=_poles(a, G, a.k, a.s[0], 'frac'), precision=a.precision,
                               progress=_progress(a, name, nm), block=_block(a), shortlist=a.shortlist,
                               **{'B': a.B, **cfg[nm]})
                res['runs'].setdefault(nm, []).append(_clean(out))
                print(name, nm, rep, f"{out['seconds']:.2f}s spmv={out['spmv']}", flush=True)
                out_dir.mkdir(parents=True, exist_ok=True)
                (out_dir / f'time_{name}{SUF}.json').write_text(json.dumps(res, indent=1))


def parser():
    P = argparse.ArgumentParser()
    P.add_argument('mode', choices=['effect', 'time'])
    P.add_argument('--graphs', nargs='+', required=True)
    P.add_argument('--N', type=int, default=None,
                   help='candidates (default: 2000 for a named graph, the configuration pool for c1..c5)')
    P.add_argument('--k', type=int, default=20)
    P.add_argument('--q', type=int, default=128)
    P.add_argument('--B', type=int, default=16)
    P.add_argument('--eps', type=float, default=1e-2)
    P.add_argument('--tol', type=float, default=1e-7)
    P.add_argument('--s', type=float, nargs='+', default=[0.25, 0.5, 0.75])
    P.add_argument('--random', type=int, default=10)
    P.add_argument('--reps', type=int, default=3)
    P.add_argument('--naive-reps', type=int, default=1)
    P.add_argument('--naive-limit', type=float, default=3600.0)
    P.add_argument('--dense-max', type=int, default=6000)
    P.add_argument('--skip-naive', action='store_true')
    P.add_argument('--cross-only', action='store_true', help='evaluate Random only under the fractional targets')
    P.add_argument('--pool', choices=['uniform', '2hop'], default='uniform')
    P.add_argument('--pool-gen', choices=['reference', 'fast'], default='reference',
                   help='uniform pool generator of a named graph (fast: vectorised, its own draw order)')
    P.add_argument('--slq', action='store_true')
    P.add_argument('--slq-probes', type=int, default=20)
    P.add_argument('--slq-steps', type=int, default=40)
    P.add_argument('--tag', default='')
    P.add_argument('--pool-seed', type=int, default=0)
    P.add_argument('--sigma', type=float, default=1.0, help='the reference module global SIGMA')
    P.add_argument('--poles', type=int, default=0, help='fixed pole count (0: composite rule)')
    P.add_argument('--kind', choices=['block', 'gauss', 'rademacher'], default='block')
    P.add_argument('--precision', choices=['fp64', 'fp32'], default='fp64')
    P.add_argument('--out', default=str(OUT))
    P.add_argument('--progress', action='store_true', help='print the bank and every round as they finish')
    P.add_argument('--only', nargs='+', default=None, help='time only these runs (Naive, B1, B)')
    P.add_argument('--block', choices=['auto', 'on', 'off'], default='auto',
                   help='blocked shifted update (auto: on for fp32, off for fp64)')
    P.add_argument('--shortlist', choices=['proxy', 'gradient'], default='proxy',
                   help='B-edge shortlist: the reference proxy or the first-order (gradient) gain')
    return P


def main(argv=None):
    global POOL, POOL_GEN, PSEED, SUF, SIGMA
    a = parser().parse_args(argv)
    POOL = a.pool
    POOL_GEN = a.pool_gen
    PSEED = a.pool_seed
    SIGMA = a.sigma
    SUF = ('' if a.pool == 'uniform' else '_2hop') + (f'_ps{a.pool_seed}' if a.pool_seed else '') + (f'_{a.tag}' if a.tag else '')
    {'effect': effect, 'time': timing}[a.mode](a)


if __name__ == '__main__':
    main()
