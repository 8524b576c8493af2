"""Autotuner of the CG iteration's launch shape, per degree profile.

Tunables, none of which changes a single result bit (every column's arithmetic and every
reduction's value are independent of them, which tests/test_gpu_batched.py checks):
  * cpl_max: columns per lane (1, 2, 4).  A row is processed by L = R/cpl lanes, so this sets the
    rows per wavefront (64/L) of the operator, residual and update kernels;
  * grid_max: block budget of the kernels that write partial sums (operator, long-row finisher,
    residual, init); more blocks means more rows in flight and more partials per column;
  * graph: replay the CG iterations as hipGraphs (fewer launch gaps; pays off for short
    iterations, i.e. small graphs or narrow blocks).
Each candidate is timed on the real operator: a fixed number of CG iterations (the solve stops at
its iteration cap) on random right-hand sides with the problem's poles and active rows.  Results
are cached by degree profile (log2 n, log2 mean degree, log2 max degree, segments), width R and
dtype, in memory and optionally in a JSON file ($POLAR_TUNE_CACHE).
"""
import json
import math
import os
import time

import numpy as np
import torch

from . import _lib as L
from . import polar_gpu as pg

_CACHE = {}
GRIDS = (256, 512, 1024, 2048, 4096, 8192)


def profile(G, S=1):
    """Degree-profile key of a graph (or union graph with S segments)."""
    deg = np.diff(G._indptr)
    mean = max(float(deg.mean()), 1.0)
    return (int(math.log2(max(G.n, 2))), round(2 * math.log2(mean)) / 2,
            int(math.log2(max(int(deg.max()), 1))), int(S))


def _key(prof, R, dtype):
    return '%s|R%d|%s' % (','.join(str(x) for x in prof), R, 'f64' if dtype == torch.float64 else 'f32')


def _load_file():
    path = os.environ.get('POLAR_TUNE_CACHE')
    if path and os.path.exists(path):
        with open(path) as f:
            _CACHE.update(json.load(f))


def _save_file():
    path = os.environ.get('POLAR_TUNE_CACHE')
    if path:
        with open(path, 'w') as f:
            json.dump(_CACHE, f, indent=1, sort_keys=True)


def _time(fn, reps):
    fn()                                   # warm (allocations, first launches)
    torch.cuda.synchronize()
    best = math.inf
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t)
    return best


def tune(op, R, tau, active, segs=None, iters=24, reps=2, log=None):
    """Best {cpl_max, grid_max, graph} for solves of width R on op (a polar_gpu.Op; with `segs`, the
    Op of a batched.UnionGraph).  tau: (p,) or (S, p) poles; active: device int32 active rows."""
    from . import batched
    _load_file()
    S = 1 if segs is None else segs.S
    key = _key(profile(op.G, S), R, op.dtype)
    if key in _CACHE:
        return dict(_CACHE[key])
    n = op.G.n
    gen = torch.Generator(device='cuda').manual_seed(1234)
    W = torch.randn((n, R), dtype=op.dtype, device='cuda', generator=gen)

    def run(cfg):
        op.cpl_max, op.grid_max = cfg['cpl_max'], cfg['grid_max']
        op._struct = None

        def go():
            try:
                if segs is None:
                    pg.msolve(op, W, tau, active, 1e-30, maxit=iters, graph=cfg['graph'])
                else:
                    batched.msolve_batched(op, segs, W, tau, active, 1e-30, maxit=iters, graph=cfg['graph'])
            except RuntimeError as e:          # the iteration cap ends every timed solve
                if 'iteration cap' not in str(e):
                    raise
        return _time(go, reps)

    esz = 8 if op.dtype == torch.float64 else 4
    cpls = [c for c in (4, 2, 1) if c * esz <= 16 and R % c == 0 and (R + c - 1) // c <= 64]
    best = dict(cpl_max=0, grid_max=0, graph=False)
    best_t = run(best)
    trials = [(dict(best), best_t)]
    for field, values in (('grid_max', GRIDS), ('cpl_max', cpls), ('graph', (True,))):
        for v in values:
            cfg = dict(best, **{field: v})
            t = run(cfg)
            trials.append((cfg, t))
            if t < best_t * 0.98:
                best, best_t = cfg, t
    op.cpl_max, op.grid_max = 0, 0
    op._struct = None
    if log:
        for cfg, t in trials:
            log(f'tune {key} {cfg} {1e3 * t / iters:.3f} ms/iter')
    best['ms_per_iter'] = 1e3 * best_t / iters
    best['default_ms_per_iter'] = 1e3 * trials[0][1] / iters
    _CACHE[key] = dict(best)
    _save_file()
    return best


def apply(op, cfg):
    """Set a tuned launch shape on op (every later solve on it uses it)."""
    op.cpl_max = int(cfg.get('cpl_max', 0))
    op.grid_max = int(cfg.get('grid_max', 0))
    op.use_graph = bool(cfg.get('graph', False))
    op._struct = None
    return op
