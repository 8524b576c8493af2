#!/usr/bin/env python
"""Benchmark of the pole-shared Krylov core on one MI355X.

Metric (k-round greedy wall time + SpMM GB/s at n=4M; FP32-mixed vs FP64 speedup):
  * step = one full POLAR greedy workload of the c2 configuration: Chung-Lu n=100,000 (gamma 2.5,
    mean degree 20, seed 1), sigma=0.05, q=64 Rademacher, 24 poles, N=1,000,000 candidates,
    k=20 rounds, B=8, for each s in {0.25, 0.5, 0.75}: the three greedy runs advance together -- the
    bank as one solve over three row segments, every re-rank solve over the three runs' shortlists
    as column groups of one operator (one adjacency pass); --schedule columns|segments|per-s for
    the alternatives, all with the same results; inputs resident in HBM.
    `value` = seconds of one step in the FP32-mixed form, timed as the reference times `seconds`
    (polar_gpu.py:196-253: after the inputs are resident, through the final sync);
    `step_wall_s` is the wall clock around the whole call including its host-side setup.
  * the plain FP64 form is timed on the same inputs (fp64_seconds; speedup = fp64/fp32) and its
    selections and marginals are compared with the FP32-mixed ones.
  * SpMM (L + sigma I) X at the c5 shape (Chung-Lu n=4,000,000, m=34.7M, R=128) in FP64 and FP32,
    GB/s on compulsory bytes; beside it the same kernel on a 2000 x 2000 lattice (same n, R), a
    locality control whose neighbour gathers hit cache.
  * roofline: the operator kernel k_opmul (the dominant kernel), its HIP-event durations from the
    library over the timed region (events around that kernel only); `kernel_profile` breaks one
    extra, untimed step down by kernel class (events around every kernel).
  * cpu_baseline: the oracle (NumPy/SciPy FP64, one core) on a bounded sample at the c2 shape,
    extrapolated to the step's column-iterations.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'k-round greedy wall time + SpMM GB/s at n=4M; FP32-mixed vs FP64 speedup'
KERNELS = ['k_opmul', 'k_longfin', 'k_alpha', 'k_resid', 'k_scalars', 'k_update']
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def traffic_of(kernel, workload):
    """Measured HBM bytes per dispatch of `kernel` on `workload` from profiles/hbm_traffic.md (rows
    `| kernel | workload | FETCH KiB | WRITE KiB | bytes |`, FETCH/WRITE scaled by the 1 GiB copy
    probe), or None when that pair has not been measured."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic.md')
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip('|').split('|')]
        if len(cells) >= 5 and cells[0] == kernel and cells[1] == workload:
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def copy_probe(torch):
    """Calibration probe for the HBM counters: one plain 1 GiB device copy (reads 1 GiB, writes
    1 GiB); its FETCH_SIZE / WRITE_SIZE readings scale the other kernels' readings."""
    src = torch.empty(1 << 28, dtype=torch.float32, device='cuda').fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    torch.cuda.empty_cache()


GRAPH = False     # direct launches: measured faster than hipGraph replay on the c2 step (0.997 vs 1.007 s)
REPLACE = False


def step_seconds(outs, union):
    """The step's timed region as the reference defines `seconds` (polar_gpu.py:196, 252-253): from
    after the inputs are resident (sketch H2D, candidate endpoint table) through the final sync;
    one batched region, or the sum over the per-s runs."""
    batched = union is not None or (SCHEDULE in ('columns', 'hybrid') and len(outs) > 1)
    return outs[0]['seconds'] if batched else float(sum(o['seconds'] for o in outs))


SCHEDULE = 'hybrid'


def run_step(pg, G, inp, sk, precision, union=None):
    """One step: the greedy for every s value.  Schedules (same results):
    'columns'  -- the s values are column groups of one operator (one adjacency pass per phase);
    'segments' -- `union` (a UnionGraph of one copy of G per s) runs them as row segments;
    'per-s'    -- one greedy run after another."""
    poles = [pg.fixed_rule(s, inp['p'], inp['sigma'], G.Lam(inp['k']))[:2] for s in inp['s_all']]
    if SCHEDULE in ('columns', 'hybrid') and len(inp['s_all']) > 1:
        from polar_amd import batched
        return batched.polar_columns(G, inp['cand'], inp['k'], list(inp['s_all']), q=inp['q'], B=inp['B'],
                                     tol=inp['tol'], sketch=sk, poles=poles, precision=precision, graph=GRAPH,
                                     replace=REPLACE and precision == 'fp32',
                                     bank_union=union if SCHEDULE == 'hybrid' else None)
    if union is not None:
        from polar_amd import batched
        S = len(inp['s_all'])
        return batched.polar_batched([G] * S, [inp['cand']] * S, inp['k'], list(inp['s_all']), q=inp['q'],
                                     B=inp['B'], tol=inp['tol'], sketch=[sk] * S, poles=poles,
                                     precision=precision, union=union, graph=GRAPH,
                                     replace=REPLACE and precision == 'fp32')
    outs = []
    for s, po in zip(inp['s_all'], poles):
        outs.append(pg.polar(G, inp['cand'], inp['k'], s, q=inp['q'], B=inp['B'], tol=inp['tol'], sketch=sk,
                             poles=po, precision=precision, graph=GRAPH, replace=REPLACE and precision == 'fp32'))
    return outs


def _time_spmm(pg, torch, lib, L, G, R, reps):
    """polar_spmm on G at width R, FP64 and FP32: ms per product (CUDA events on the launch
    stream) and GB/s on compulsory bytes (CSR once, X read once, Y written once)."""
    n = G.n
    res = {}
    for prec, dt, esz in (('fp64', torch.float64, 8), ('fp32', torch.float32, 4)):
        op = pg.Op(G, precision=prec)
        X = torch.randn((n, R), dtype=dt, device='cuda')
        Y = torch.empty_like(X)
        s = op.c_op(dt)
        ws = torch.empty(max(lib.polar_spmm_ws_bytes(s, R), 256), dtype=torch.uint8, device='cuda')
        st = torch.cuda.current_stream()

        def go():
            L.check(lib.polar_spmm(s, 0.0, L.ptr(X), R, L.ptr(Y), R, R, L.ptr(ws), ws.numel(), L.stream_ptr()))
        go()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            go()
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nnz = G.nnz
        ours = 4 * (n + 1) + 4 * nnz + esz * n + 2 * n * R * esz            # value-free format
        refb = 4 * (n + 1) + (n + nnz) * (4 + esz) + 2 * n * R * esz        # CSR with diagonal
        res[prec] = dict(ms=ms, gbps=ours / ms / 1e6, bytes=ours, gbps_ref_format=refb / ms / 1e6,
                         frac=ours / ms / 1e6 / HBM_PEAK_GBPS)
        del X, Y, ws
        torch.cuda.empty_cache()
    return res


def spmm_bench(pg, torch, lib, L, n_scale, reps=10):
    from polar_amd import configs, graphs
    spec = configs.SPECS['c5']
    n = int(spec['n'] * n_scale)
    t = time.time()
    e = graphs.chung_lu(n, spec['gamma'], spec['mean_deg'], spec['seed'], m=int(spec['m'] * n_scale))
    log(f'c5 graph n={n} m={len(e)} built in {time.time() - t:.1f}s')
    G = pg.Graph(n, e, spec['sigma'])
    R = 128
    res = _time_spmm(pg, torch, lib, L, G, R, reps)
    return dict(n=n, m=len(e), nnz_offdiag=G.nnz, R=R, dmax=G.dmax, **res)


def spmm_lattice_bench(pg, torch, lib, L, n_scale, reps=10):
    """Locality control for the SpMM: a 2-D lattice with the c5 row count (4-neighbour, side
    sqrt(n)), whose neighbour rows i +- 1 and i +- side stay cache-resident.  Same kernel, same
    R; it shows what the kernel reaches when the gathers hit cache."""
    side = int(round(math.sqrt(4_000_000 * n_scale)))
    n = side * side
    idx = np.arange(n, dtype=np.int64).reshape(side, side)
    e = np.concatenate([np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1),
                        np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], 1)])
    G = pg.Graph(n, e, 0.02)
    R = 128
    res = _time_spmm(pg, torch, lib, L, G, R, reps)
    return dict(n=n, m=len(e), nnz_offdiag=G.nnz, R=R, graph=f'{side}x{side} lattice', **res)


def cpu_baseline(inp, step_colits, seconds_budget=20.0):
    """Oracle msolve at the c2 shape on an 8-column block, bounded by iteration count; the cost
    per column-iteration is scaled by the GPU step's column-iterations."""
    from oracle import polar_oracle as O
    Go = O.Graph(inp['n'], inp['edges'], inp['sigma'])
    Z, F, _ = O.sketches(Go, inp['cand'][:16], inp['q'], inp['seed'], kind='rademacher')
    active = np.unique(inp['cand'])
    tau, om, _ = O.fixed_rule(0.5, inp['p'], inp['sigma'], Go.Lam(inp['k']))
    W = Z[:, :8]
    its = 8
    t0 = time.perf_counter()
    while True:
        try:
            O.msolve(O.Op(Go), W, tau, active, inp['tol'], maxit=its)
        except RuntimeError:
            pass
        el = time.perf_counter() - t0
        if el > seconds_budget / 4 or its >= 4096:
            break
        its *= 2
        t0 = time.perf_counter()
    per_colit = el / (its * W.shape[1])
    return dict(value=per_colit * step_colits, unit='s', cores=1, kind='port',
                sample=f'oracle msolve on 8 columns x {its} CG iterations at the c2 shape '
                       f'(n={inp["n"]}, p={inp["p"]}, nA={len(active)}): {per_colit * 1e3:.3f} ms per '
                       f'column-iteration, scaled by the step\'s {step_colits} column-iterations',
                cpu_model=_cpu_model(), os_cpu_count=os.cpu_count(),
                affinity=len(os.sched_getaffinity(0)), blas_threads=_blas_threads(),
                note='scipy.sparse products are single-threaded; the oracle runs on one core')


def cpu_c1(pg):
    """The CPU-runnable c1 workload run to completion by the oracle (NumPy/SciPy FP64) and by the
    FP64 HIP path on the same inputs: BA n=2,000 (m0=5, seed 0), sigma=0.1, s=0.5, q=32
    Rademacher, 16 poles, N=5,000, k=10, B=1."""
    from oracle import polar_oracle as O
    from polar_amd import configs
    c = configs.make('c1')
    Go = O.Graph(c['n'], c['edges'], c['sigma'])
    sk = O.sketches(Go, c['cand'], c['q'], c['seed'], kind='rademacher')
    kw = dict(q=c['q'], B=c['B'], tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']))
    t = time.perf_counter()
    ro = O.polar(Go, c['cand'], c['k'], c['s'], **kw)
    cpu_s = time.perf_counter() - t
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    pg.polar(G, c['cand'], c['k'], c['s'], **kw)                    # warm
    rg = pg.polar(G, c['cand'], c['k'], c['s'], **kw)
    return dict(workload='c1: BA n=2000 m0=5, sigma=0.1, s=0.5, q=32, p=16, N=5000, k=10, B=1, FP64',
                cpu_seconds=cpu_s, gpu_seconds=rg['seconds'], cpu_over_gpu=cpu_s / rg['seconds'],
                selections_identical=ro['sequence'] == rg['sequence'], cores=1)


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return {i.get('internal_api', i.get('user_api', '?')): i.get('num_threads') for i in threadpool_info()}
    except Exception:
        return None


def _cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='c2')
    ap.add_argument('--scale', type=float, default=1.0)
    ap.add_argument('--skip-fp64', action='store_true')
    ap.add_argument('--skip-cpu', action='store_true')
    ap.add_argument('--skip-spmm', action='store_true')
    ap.add_argument('--spmm-scale', type=float, default=1.0)
    ap.add_argument('--schedule', choices=['hybrid', 'columns', 'segments', 'per-s'], default='hybrid',
                    help='how the s values of a step share the GPU (same results)')
    ap.add_argument('--graph', choices=['on', 'off'], default='off',
                    help='hipGraph replay of the CG iterations of unprofiled solves (the FP64 leg)')
    ap.add_argument('--replace', action='store_true', help='FP64 residual replacement in the FP32-mixed step')
    ap.add_argument('--spmm-only', action='store_true',
                    help='only the n=4M SpMM leg and the 1 GiB copy probe (for counter runs)')
    a = ap.parse_args()
    global GRAPH, REPLACE, SCHEDULE
    SCHEDULE = a.schedule
    GRAPH = a.graph == 'on'
    REPLACE = a.replace
    if a.gpus != 1:
        print(f'bench.py measures one GPU only (asked for --gpus {a.gpus}); the pole-sharded multi-GPU '
              'layout is a design note, not built.', file=sys.stderr)
        sys.exit(2)

    import torch
    from polar_amd import _lib as L
    from polar_amd import configs
    from polar_amd import polar_gpu as pg
    lib = L.require_gpu()
    torch.cuda.set_device(0)

    if a.spmm_only:
        copy_probe(torch)
        print(json.dumps(dict(spmm_n4m=spmm_bench(pg, torch, lib, L, a.spmm_scale),
                              spmm_lattice_n4m=spmm_lattice_bench(pg, torch, lib, L, a.spmm_scale))), flush=True)
        return

    t = time.time()
    inp = configs.make(a.config, a.scale)
    G = pg.Graph(inp['n'], inp['edges'], inp['sigma'])
    sk = pg.sketches(G, inp['cand'], inp['q'], inp['seed'], kind='rademacher', on_device='lazy')
    torch.cuda.synchronize()
    log(f'{a.config}: n={inp["n"]} m={inp["m"]} N={inp["N"]} nA={len(np.unique(inp["cand"]))} '
        f'dmax={inp["dmax"]} inputs in {time.time() - t:.1f}s')

    copy_probe(torch)
    union = None
    if a.schedule in ('segments', 'hybrid') and len(inp['s_all']) > 1:
        from polar_amd import batched
        union = batched.UnionGraph([G] * len(inp['s_all']))

    for w in range(a.warmup):
        run_step(pg, G, inp, sk, 'fp32', union)
        log(f'warmup {w} done')
    def read_prof():
        out = {}
        for cls, name in enumerate(KERNELS + ['solves', 'iterations']):
            ms, la, by = (np.zeros(1), np.zeros(1, np.int64), np.zeros(1))
            lib.polar_prof_get(cls, ms.ctypes.data_as(L.vp), la.ctypes.data_as(L.vp), by.ctypes.data_as(L.vp))
            out[name] = dict(ms=float(ms[0]), launches=int(la[0]), bytes=float(by[0]))
        return out

    # timed steps: HIP events around the operator kernel on one CG iteration in every 8 (profiler
    # mode 3; the sampled position rotates), so the roofline kernel is timed live over the timed
    # region while the step itself is barely perturbed
    L.check(lib.polar_prof_set(3), 'prof_set')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs32 = None
    timed = 0.0
    for i in range(a.steps):
        outs32 = run_step(pg, G, inp, sk, 'fp32', union)
        timed += step_seconds(outs32, union)
        log(f'step {i} done')
    torch.cuda.synchronize()
    wall32 = (time.perf_counter() - t0) / a.steps
    el32 = timed / a.steps
    prof_t = read_prof()
    op_prof, op_total = prof_t['k_opmul'], prof_t['iterations']['launches']
    # one more, untimed step with events around every kernel class (the per-class breakdown)
    lib.polar_prof_set(1)
    run_step(pg, G, inp, sk, 'fp32', union)
    torch.cuda.synchronize()
    prof = read_prof()
    lib.polar_prof_set(0)
    log('fp32 step %.3fs' % el32, json.dumps(prof))

    el64 = None
    agree = None
    if not a.skip_fp64:
        run_step(pg, G, inp, sk, 'fp64', union) if a.warmup else None
        torch.cuda.synchronize()
        outs64 = run_step(pg, G, inp, sk, 'fp64', union)
        torch.cuda.synchronize()
        el64 = step_seconds(outs64, union)
        same = all(o32['sequence'] == o64['sequence'] for o32, o64 in zip(outs32, outs64))
        mrel = max(float(np.max(np.abs(np.array(o32['marginals']) - o64['marginals']) / np.abs(o64['marginals'])))
                   for o32, o64 in zip(outs32, outs64))
        trel = max(float(abs(o32['trace_estimate'][0] - o64['trace_estimate'][0]) / abs(o64['trace_estimate'][0]))
                   for o32, o64 in zip(outs32, outs64))
        agree = dict(selections_identical=same, max_rel_marginal_err=mrel, max_rel_trace_err=trel)
        log('fp64 step %.3fs' % el64, json.dumps(agree))

    spmm = lattice = None
    if not a.skip_spmm:
        spmm = spmm_bench(pg, torch, lib, L, a.spmm_scale)
        log('spmm', json.dumps(spmm))
        lattice = spmm_lattice_bench(pg, torch, lib, L, a.spmm_scale)
        log('spmm lattice', json.dumps(lattice))

    colits = int(sum(o['spmv'] for o in outs32))
    dom = max(KERNELS, key=lambda k: prof[k]['ms'])       # k_opmul on every config measured
    if dom != 'k_opmul':
        log(f'note: the dominant kernel class of the profiled step is {dom}; the roofline is k_opmul')
    dom = 'k_opmul'
    d = op_prof
    achieved = d['bytes'] / (d['ms'] * 1e-3) / 1e9 if d['ms'] > 0 else 0.0
    wl = f'{a.config}-fp32' + ('' if a.scale == 1.0 else f'-x{a.scale}')
    roofline = dict(kernel=dom, bound='hbm', achieved=achieved, peak=HBM_PEAK_GBPS, unit='GB/s',
                    frac=achieved / HBM_PEAK_GBPS, traffic=traffic_of(dom, wl),
                    avg_launch_us=d['ms'] * 1e3 / max(d['launches'], 1),
                    algorithmic_bytes_per_launch=d['bytes'] / max(d['launches'], 1),
                    timed_launches=d['launches'], launches=op_total,
                    share_of_timed_region=d['ms'] / max(d['launches'], 1) * op_total * 1e-3 / (el32 * a.steps))
    cpu = None
    if not a.skip_cpu:
        cpu = cpu_baseline(inp, colits)
        cpu['c1_to_completion'] = cpu_c1(pg)
        log('cpu', json.dumps(cpu))
    out = dict(metric=METRIC, value=el32, unit='s', n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=el32 * 1e3, step_wall_s=wall32, higher_is_better=False, scaling='weak', vs_baseline=None,
               dtype='fp32 vectors / fp64 recurrences', data='synthetic',
               config=dict(workload=f'{a.config}: Chung-Lu n={inp["n"]}, m={inp["m"]}, sigma={inp["sigma"]}, '
                                    f's={list(inp["s_all"])}, q={inp["q"]}, p={inp["p"]}, N={inp["N"]}, '
                                    f'k={inp["k"]}, B={inp["B"]}, tol={inp["tol"]}',
                           precision='fp32-mixed', column_iterations=colits,
                           hipgraph=GRAPH, residual_replacement=REPLACE,
                           residual_replacement_note=('FP64 residual replacement every 24 iterations'
                                                      if REPLACE else 'off by default: --replace turns on FP64 '
                                                      'residual replacement every 24 iterations (measured: same '
                                                      'FP32 error, 33% slower step)'),
                           s_schedule={'hybrid': f'{len(inp["s_all"])} s values: bank as row segments, re-rank as column groups',
                                       'columns': f'{len(inp["s_all"])} s values as column groups of one operator',
                                       'segments': f'{len(inp["s_all"])} s values as segments of one batched solve',
                                       'per-s': 'one greedy run per s value'}[a.schedule],
                           rhs=int(sum(o['rhs'] for o in outs32))),
               fp64_seconds=el64, fp32_vs_fp64_speedup=(el64 / el32) if el64 else None,
               fp32_vs_fp64=agree, spmm_n4m=spmm, spmm_lattice_n4m=lattice, roofline=roofline, cpu_baseline=cpu,
               kernel_profile=prof, kernel_profile_note='one extra untimed FP32 step with HIP events around '
               'every kernel class (the timed steps carry events around k_opmul on one iteration in 8)')
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
