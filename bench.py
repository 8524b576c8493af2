#!/usr/bin/env python
"""Benchmark of the pole-shared Krylov core on one MI355X.

Metric (k-round greedy wall time + SpMM GB/s at n=4M; FP32-mixed vs FP64 speedup):
  * step = one full POLAR greedy workload of the c2 configuration: Chung-Lu n=100,000 (gamma 2.5,
    mean degree 20, seed 1), sigma=0.05, q=64 Rademacher, 24 poles, N=1,000,000 candidates,
    k=20 rounds, B=8, for each s in {0.25, 0.5, 0.75}, in the FP32-mixed form (FP32 vectors and
    SpMM, FP64 recurrences and reductions, FP64 residual replacement every 24 iterations): the three
    greedy runs advance together -- the bank as one solve over three row segments, every re-rank
    solve over the three runs' shortlists as column groups of one operator; inputs resident in HBM.
    `value` = seconds of one step, timed as the reference times `seconds` (polar_gpu.py:196-253:
    after the inputs are resident, through the final sync); K steps are timed after W warmups.
  * beside it on the same inputs: the FP32-mixed step without residual replacement (its cost), the
    plain FP64 step (pinned per-iteration shifted update) and the FP64 step with the blocked
    update (so the FP32 speedup can be read for precision alone), with the FP32-vs-FP64 agreement.
  * c5_greedy: the headline configuration -- Chung-Lu n=4,000,000, m=34.7M, sigma=0.02, s=0.5,
    q=128, 32 poles, N=10,000,000, k=50, B=16 -- run once in the FP32-mixed form and once in each
    FP64 form: seconds, column-iterations, speedups, shared selections, marginal and trace errors,
    and a per-round record of the picks that differ (isolated-pair picks, whose exact gains are
    equal, with the FP64 re-rank score gap).
  * SpMM (L + sigma I) X on the c5 graph at R=128 in FP64 and FP32, GB/s on compulsory bytes, and
    the same kernel on a 2000 x 2000 lattice (same n, R): a locality control.
  * roofline: the operator kernel k_opmul of the c2 step (the dominant kernel), its HIP-event
    durations from the library over the timed region (events around that kernel only);
    `kernel_profile` breaks one extra, untimed step down by kernel class.
  * cpu_baseline: the oracle (NumPy/SciPy FP64) polar() on the c2 inputs for s=0.5, its block
    solves' columns split over worker processes, under the reference's time_limit convention
    (polar_gpu.py:215-216), run as a child process on the host while the GPU runs the c5 legs;
    plus the c1 workload run to completion on one core beside the FP64 HIP path.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--skip-c5] [--skip-cpu] ...
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'k-round greedy wall time + SpMM GB/s at n=4M; FP32-mixed vs FP64 speedup'
KERNELS = ['k_opmul', 'k_longfin', 'k_alpha', 'k_resid', 'k_scalars', 'k_update']
HBM_PEAK_GBPS = 8000.0
T_START = time.time()


def log(*a):
    print('[bench %5.0fs]' % (time.time() - T_START), *a, file=sys.stderr, flush=True)


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
REPLACE = None    # FP64 residual replacement: None = the library default (on for FP32-mixed)


def step_work(outs):
    """Column-iterations the step performed: a shared bank solve counts once, not once per s."""
    tot = sum(o['spmv'] for o in outs)
    if outs and outs[0].get('shared_bank'):
        tot -= (len(outs) - 1) * outs[0]['bank_spmv']
    return int(tot)


def step_seconds(outs, union):
    """The step's timed region as the reference defines `seconds` (polar_gpu.py:196, 252-253): from
    after the inputs are resident (sketch H2D, candidate endpoint table) through the final sync;
    one batched region, or the sum over the per-s runs."""
    batched = union is not None or (SCHEDULE in ('columns', 'hybrid') and len(outs) > 1)
    return outs[0]['seconds'] if batched else float(sum(o['seconds'] for o in outs))


SCHEDULE = 'hybrid'


def run_step(pg, G, inp, sk, precision, union=None, replace=None, block=None, shared_bank=None):
    """One step: the greedy for every s value.  Schedules (same results):
    'columns'  -- the s values are column groups of one operator (one adjacency pass per phase);
    'segments' -- `union` (a UnionGraph of one copy of G per s) runs them as row segments;
    'per-s'    -- one greedy run after another."""
    poles = [pg.fixed_rule(s, inp['p'], inp['sigma'], G.Lam(inp['k']))[:2] for s in inp['s_all']]
    if SCHEDULE in ('columns', 'hybrid') and len(inp['s_all']) > 1:
        from polar_amd import batched
        return batched.polar_columns(G, inp['cand'], inp['k'], list(inp['s_all']), q=inp['q'], B=inp['B'],
                                     tol=inp['tol'], sketch=sk, poles=poles, precision=precision, graph=GRAPH,
                                     replace=replace, block=block, shared_bank=shared_bank,
                                     bank_union=union if SCHEDULE == 'hybrid' else None)
    if union is not None:
        from polar_amd import batched
        S = len(inp['s_all'])
        return batched.polar_batched([G] * S, [inp['cand']] * S, inp['k'], list(inp['s_all']), q=inp['q'],
                                     B=inp['B'], tol=inp['tol'], sketch=[sk] * S, poles=poles,
                                     precision=precision, union=union, graph=GRAPH,
                                     replace=replace, block=block)
    outs = []
    for s, po in zip(inp['s_all'], poles):
        outs.append(pg.polar(G, inp['cand'], inp['k'], s, q=inp['q'], B=inp['B'], tol=inp['tol'], sketch=sk,
                             poles=po, precision=precision, graph=GRAPH, replace=replace, block=block))
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


def spmm_bench(pg, torch, lib, L, G, reps=10):
    """polar_spmm on the c5 graph (Chung-Lu n=4M, m=34.7M) at R=128, FP64 and FP32."""
    R = 128
    res = _time_spmm(pg, torch, lib, L, G, R, reps)
    return dict(n=G.n, m=G.m, nnz_offdiag=G.nnz, R=R, dmax=G.dmax, **res)


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


def flush_bench(pg, torch, graphs):
    """The window flush alone at the n=4M bank's width: an FP32 blocked solve on a Chung-Lu graph
    (n=200k, mean degree 20), 128 Rademacher columns, 32 poles, every row active and a tolerance
    nothing reaches, so every (pole, column) pair is changed in each of the 8 windows of the 64
    iterations it runs (the solve then stops at its iteration cap).  The flush's duration is read
    from a kernel trace; its bytes per launch are the compulsory ones: P and X of every pole read
    and written, the 8 window residuals read."""
    n, R, p, iters = 200_000, 128, 32, 64
    G = pg.Graph(n, graphs.chung_lu(n, 2.5, 20, 1), 0.05)
    W = torch.from_numpy(np.random.default_rng(0).choice([-1.0, 1.0], size=(n, R))).float().cuda()
    tau = np.geomspace(0.05, 400.0, p)
    for replace in (False, True):
        try:
            pg.msolve(pg.Op(G, precision='fp32'), W, tau, np.arange(n), 1e-30, maxit=iters, block=True,
                      replace=replace)
        except RuntimeError as ex:
            if 'iteration cap' not in str(ex):
                raise
    torch.cuda.synchronize()
    return dict(n=n, R=R, poles=p, iterations=iters, windows=iters // 8, runs=['replace off', 'replace on'],
                flush_bytes=p * n * R * 4 * 4 + 8 * n * R * 4,
                seed_bytes_with_replace=n * R * (4 + 4 + 4 + 8 + 8))


def true_marginals_bench(pg, torch, configs, k=20, reps=2):
    """The certified evaluator (polar_gpu.true_marginals, eps=1e-9, tol=1e-10, FP64) on the c2
    graph along a random k-edge sequence of its pool: the positions as column groups of one solve
    against the sequential loop (together=1); best of `reps` runs each."""
    c = configs.make('c2', 1.0)
    G = pg.Graph(c['n'], c['edges'], c['sigma'])
    seq = pg.random_set(c['cand'], k, 1)
    out = {}
    vals = {}
    for together in (1, 64):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            g, cert = pg.true_marginals(G, c['cand'], seq, 0.5, together=together)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        vals[together] = g
        out[f'together_{together}'] = dict(seconds=best, spmv=cert['spmv'], p=cert['p'])
    out.update(k=k, n=G.n, max_rel_diff=float(np.max(np.abs(vals[64] - vals[1]) / np.abs(vals[1]))))
    return out


def _rel_max(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    n = min(len(a), len(b))
    if n == 0:
        return None
    return float(np.max(np.abs(a[:n] - b[:n]) / np.maximum(np.abs(b[:n]), 1e-300)))


def agreement(o32, o64):
    """FP32-mixed against FP64 on the same inputs: selections, marginals, trace estimates."""
    same = all(a['sequence'] == b['sequence'] for a, b in zip(o32, o64))
    shared = [len(set(a['sequence']) & set(b['sequence'])) for a, b in zip(o32, o64)]
    return dict(selections_identical=same, shared_selections=shared,
                max_rel_marginal_err=max(_rel_max(a['marginals'], b['marginals']) for a, b in zip(o32, o64)),
                max_rel_trace_err=max(_rel_max(a['trace_estimate'], b['trace_estimate']) for a, b in zip(o32, o64)))


def tie_record(c, deg, runs, ref='fp64', other='fp32'):
    """Rounds where two runs pick differently.  A pick (a, b) joining two vertices of degree 0 in
    that run's current graph (base edges plus its earlier picks) has the exact fractional gain
    sigma^-s - (sigma+2)^-s, whatever the pair; for such rounds the two picks' exact gains are equal
    and the choice is a tie.  Also the re-rank scores of the reference run's shortlist: the top-2
    relative gap and how many shortlisted pairs lie within 1e-6 of the best."""
    cand = c['cand']
    sig, s = c['sigma'], c['s']
    closed = sig ** (-s) - (sig + 2.0) ** (-s)

    def isolated_flags(seq):
        used = set()
        out = []
        for e in seq:
            a, b = (int(x) for x in cand[e])
            out.append(bool(deg[a] == 0 and deg[b] == 0 and a not in used and b not in used))
            used.update((a, b))
        return out

    ra, rb = runs[ref], runs[other]
    ia, ib = isolated_flags(ra['sequence']), isolated_flags(rb['sequence'])
    rows = []
    for r, (ea, eb) in enumerate(zip(ra['sequence'], rb['sequence'])):
        if ea == eb:
            continue
        sc = np.asarray(ra['trace_rounds'][r]['scores'], float)
        best = float(sc.max()) if len(sc) else float('nan')
        rows.append(dict(round=r, **{f'{ref}_pick': ea, f'{other}_pick': eb,
                                     f'{ref}_isolated_pair': ia[r], f'{other}_isolated_pair': ib[r],
                                     f'{ref}_marginal': ra['marginals'][r], f'{other}_marginal': rb['marginals'][r],
                                     f'{ref}_rerank_gap': ra['rerank_gap'][r], f'{other}_rerank_gap': rb['rerank_gap'][r],
                                     f'{ref}_near_ties_1e-6': int(np.sum(np.abs(sc - best) <= 1e-6 * abs(best)))}))
    return dict(closed_form_isolated_pair_gain=closed, rounds_differing=len(rows),
                all_differing_picks_isolated_pairs=all(x[f'{ref}_isolated_pair'] and x[f'{other}_isolated_pair']
                                                       for x in rows),
                max_ref_rerank_gap_differing=max((x[f'{ref}_rerank_gap'] for x in rows), default=None),
                isolated_pair_picks={ref: int(sum(ia)), other: int(sum(ib))},
                differing=rows)


def c5_greedy(pg, torch, c, G, k, legs):
    """The headline configuration, once per form, on the same graph, pool, sketch and poles."""
    sk = pg.sketches(G, c['cand'], c['q'], c['seed'], kind='rademacher', on_device='lazy')
    kw = dict(q=c['q'], B=c['B'], tol=c['tol'], sketch=sk, poles=(c['tau'], c['om']))
    forms = dict(fp32=dict(precision='fp32'), fp64=dict(precision='fp64'),
                 fp64_blocked=dict(precision='fp64', block=True))
    runs = {}
    for name in legs:
        tr = []
        t = time.time()

        def prog(stage, info, name=name):
            if stage == 'bank' or info.get('round', 0) % 10 == 9:
                log(f'c5 {name} {stage} {json.dumps({a: info[a] for a in info if a != "marginal"})}')
        r = pg.polar(G, c['cand'], k, c['s'], trace=tr, progress=prog, **forms[name], **kw)
        r['trace_rounds'] = tr
        r['wall_s'] = time.time() - t
        runs[name] = r
        log(f'c5 {name}: {r["seconds"]:.1f} s, spmv {r["spmv"]}')
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    out = dict(workload=f'c5: Chung-Lu n={c["n"]}, gamma 2.2, m={c["m"]}, seed {c["seed"]}; sigma={c["sigma"]}, '
                        f's={c["s"]}, q={c["q"]} Rademacher, p={c["p"]} poles, N={c["N"]}, k={k}, B={c["B"]}, '
                        f'tol={c["tol"]}',
               legs={n: dict(seconds=r['seconds'], wall_s=r['wall_s'], column_iterations=r['spmv'], rhs=r['rhs'],
                             replace=r['replace'], rounds=r['rounds'],
                             form={'fp32': 'FP32-mixed (blocked update, FP64 residual replacement every 24 '
                                           'iterations)', 'fp64': 'FP64, pinned per-iteration shifted update',
                                   'fp64_blocked': 'FP64, blocked shifted update'}[n])
                     for n, r in runs.items()})
    if 'fp32' in runs:
        for ref in ('fp64', 'fp64_blocked'):
            if ref in runs:
                out[f'speedup_vs_{ref}'] = runs[ref]['seconds'] / runs['fp32']['seconds']
        if 'fp64' in runs:
            deg = np.bincount(c['edges'][:, 0], minlength=c['n']) + np.bincount(c['edges'][:, 1], minlength=c['n'])
            a, b = runs['fp32'], runs['fp64']
            first = next((r for r, (x, y) in enumerate(zip(a['sequence'], b['sequence'])) if x != y), None)
            out['fp32_vs_fp64'] = dict(agreement([a], [b]), first_differing_round=first,
                                       marginals_fp32=a['marginals'], marginals_fp64=b['marginals'])
            out['ties'] = tie_record(c, deg, runs)
    if 'fp64' in runs and 'fp64_blocked' in runs:
        out['fp64_blocked_vs_fp64'] = agreement([runs['fp64_blocked']], [runs['fp64']])
    return out


def start_cpu_child(a):
    """The CPU-baseline child process, started before this process touches the GPU; it prepares
    its inputs, then waits for a line on stdin before it starts its timed run."""
    cmd = [sys.executable, '-m', 'oracle.cpu_baseline', '--config', 'c2', '--s', '0.5', '--workers',
           str(a.cpu_workers), '--time-limit', str(a.cpu_time_limit), '--wait']
    return subprocess.Popen(cmd, cwd=ROOT, stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                            text=True)


def finish_cpu_child(proc, deadline):
    try:
        out, err = proc.communicate(timeout=max(1.0, deadline - time.time()))
    except subprocess.TimeoutExpired:
        proc.kill()
        proc.communicate()
        return dict(error=f'not finished by the bench deadline ({deadline - T_START:.0f} s after start)')
    lines = [x for x in out.splitlines() if x.startswith('{')]
    if proc.returncode != 0 or not lines:
        return dict(error=f'exit {proc.returncode}: {err[-500:]}')
    return json.loads(lines[-1])


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
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='c2')
    ap.add_argument('--scale', type=float, default=1.0)
    ap.add_argument('--skip-fp64', action='store_true')
    ap.add_argument('--skip-cpu', action='store_true')
    ap.add_argument('--skip-spmm', action='store_true')
    ap.add_argument('--skip-c5', action='store_true')
    ap.add_argument('--c5-k', type=int, default=50)
    ap.add_argument('--c5-legs', default='fp32,fp64,fp64_blocked')
    ap.add_argument('--cpu-workers', type=int, default=12)
    ap.add_argument('--cpu-time-limit', type=float, default=60.0)
    ap.add_argument('--deadline', type=float, default=540.0,
                    help='seconds after start by which the CPU-baseline child must have finished')
    ap.add_argument('--schedule', choices=['hybrid', 'columns', 'segments', 'per-s'], default='hybrid',
                    help='how the s values of a step share the GPU (same results)')
    ap.add_argument('--graph', choices=['on', 'off'], default='off',
                    help='hipGraph replay of the CG iterations')
    ap.add_argument('--replace', choices=['default', 'on', 'off'], default='default',
                    help='FP64 residual replacement in the FP32-mixed step (default: on)')
    ap.add_argument('--spmm-only', action='store_true',
                    help='only the n=4M SpMM leg and the 1 GiB copy probe (for counter runs)')
    ap.add_argument('--flush-only', action='store_true',
                    help='only the window-flush solve (every pole changed; for kernel traces)')
    ap.add_argument('--tm-only', action='store_true',
                    help='only the certified-marginals evaluator on the c2 graph (positions together '
                         'against one at a time)')
    a = ap.parse_args()
    global GRAPH, REPLACE, SCHEDULE
    SCHEDULE = a.schedule
    GRAPH = a.graph == 'on'
    REPLACE = {'default': None, 'on': True, 'off': False}[a.replace]
    if a.gpus != 1:
        print(f'bench.py measures one GPU only (asked for --gpus {a.gpus}); the pole-sharded multi-GPU '
              'layout is a design note, not built.', file=sys.stderr)
        sys.exit(2)
    cpu_proc = None if (a.skip_cpu or a.spmm_only or a.flush_only or a.tm_only) else start_cpu_child(a)

    import torch
    from polar_amd import _lib as L
    from polar_amd import configs
    from polar_amd import polar_gpu as pg
    lib = L.require_gpu()
    torch.cuda.set_device(0)

    if a.spmm_only:
        copy_probe(torch)
        c = configs.make('c5')
        G5 = pg.Graph(c['n'], c['edges'], c['sigma'])
        print(json.dumps(dict(spmm_n4m=spmm_bench(pg, torch, lib, L, G5),
                              spmm_lattice_n4m=spmm_lattice_bench(pg, torch, lib, L, 1.0))), flush=True)
        return
    if a.flush_only:
        from polar_amd import graphs
        print(json.dumps(dict(flush=flush_bench(pg, torch, graphs))), flush=True)
        return
    if a.tm_only:
        print(json.dumps(dict(true_marginals=true_marginals_bench(pg, torch, configs))), flush=True)
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
        run_step(pg, G, inp, sk, 'fp32', union, replace=REPLACE)
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
        outs32 = run_step(pg, G, inp, sk, 'fp32', union, replace=REPLACE)
        timed += step_seconds(outs32, union)
        log(f'step {i} done')
    torch.cuda.synchronize()
    wall32 = (time.perf_counter() - t0) / a.steps
    el32 = timed / a.steps
    prof_t = read_prof()
    op_prof, op_total = prof_t['k_opmul'], prof_t['iterations']['launches']
    # one more, untimed step with events around every kernel class (the per-class breakdown)
    lib.polar_prof_set(1)
    run_step(pg, G, inp, sk, 'fp32', union, replace=REPLACE)
    torch.cuda.synchronize()
    prof = read_prof()
    lib.polar_prof_set(0)
    replaced = bool(outs32[0]['replace'])
    log('fp32 step %.3fs (replacement %s)' % (el32, replaced), json.dumps(prof))

    # the same step with the other replacement setting: the cost of residual replacement
    other = run_step(pg, G, inp, sk, 'fp32', union, replace=not replaced)
    torch.cuda.synchronize()
    el_other = step_seconds(other, union)
    repl_cost = dict(step_with_replacement_s=el32 if replaced else el_other,
                     step_without_replacement_s=el_other if replaced else el32)
    repl_cost['ratio'] = repl_cost['step_with_replacement_s'] / repl_cost['step_without_replacement_s']
    log('replacement cost', json.dumps(repl_cost))

    fp64 = {}
    if not a.skip_fp64:
        # the plain FP64 path (pinned per-iteration update, per-s seeds: bit-identical to the oracle), and
        # FP64 on the FP32 leg's schedule (blocked update, shared-seed bank): the precision alone
        for name, blk, shb in (('fp64', None, None), ('fp64_blocked', True, True)):
            if a.warmup:
                run_step(pg, G, inp, sk, 'fp64', union, block=blk, shared_bank=shb)
            torch.cuda.synchronize()
            o64 = run_step(pg, G, inp, sk, 'fp64', union, block=blk, shared_bank=shb)
            torch.cuda.synchronize()
            fp64[name] = dict(seconds=step_seconds(o64, union), outs=o64)
            log(f'{name} step %.3fs' % fp64[name]['seconds'])

    if cpu_proc is not None:
        cpu_proc.stdin.write('go\n')
        cpu_proc.stdin.flush()
        log('cpu baseline child started its timed run')

    spmm = lattice = c5 = None
    if not (a.skip_spmm and a.skip_c5):
        t = time.time()
        c = configs.make('c5')
        G5 = pg.Graph(c['n'], c['edges'], c['sigma'])
        log(f'c5 inputs (graph, pool) in {time.time() - t:.1f}s')
        if not a.skip_spmm:
            spmm = spmm_bench(pg, torch, lib, L, G5)
            log('spmm', json.dumps(spmm))
            lattice = spmm_lattice_bench(pg, torch, lib, L, 1.0)
            log('spmm lattice', json.dumps(lattice))
        if not a.skip_c5:
            try:
                c5 = c5_greedy(pg, torch, c, G5, a.c5_k, [x for x in a.c5_legs.split(',') if x])
            except Exception as ex:          # the c2 record still prints
                c5 = dict(error=repr(ex))
                log('c5 failed', repr(ex))
        del G5, c
        torch.cuda.empty_cache()

    colits = step_work(outs32)
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
    if cpu_proc is not None:
        run = finish_cpu_child(cpu_proc, T_START + a.deadline)
        log('cpu child', json.dumps({x: run[x] for x in run if x != 'sequence'}))
        gpu_s05 = outs32[len(outs32) // 2] if len(outs32) > 1 else outs32[0]
        if 'error' in run:
            cpu = dict(value=None, unit='s', cores=a.cpu_workers, kind='port', error=run['error'])
        else:
            same = run['sequence'] == gpu_s05['sequence'][:len(run['sequence'])]
            cpu = dict(value=run['seconds'], unit='s', cores=run['workers'], kind='port',
                       sample=f'oracle polar() (NumPy/SciPy FP64) on the c2 inputs for s=0.5 (one of the step\'s '
                              f'three s values), time_limit={run["time_limit"]:.0f} s in the reference\'s convention: '
                              f'>{run["seconds"]:.0f} s ({run["rounds"]} of {run["k"]} rounds); bank (all '
                              f'{2 * run["q"]} columns) {run["bank_seconds"]:.1f} s, then {run["per_round_seconds"]:.1f} s '
                              f'per round; block-solve columns split over {run["workers"]} worker processes',
                       rounds=run['rounds'], of_rounds=run['k'], truncated=run['truncated'],
                       bank_seconds=run['bank_seconds'], per_round_seconds=run['per_round_seconds'],
                       column_iterations=run['spmv'], sequence_prefix_matches_gpu=same,
                       cpu_model=run['cpu_model'], os_cpu_count=run['os_cpu_count'], affinity=run['affinity'],
                       blas_threads=_blas_threads(),
                       note='ran on the host concurrently with the GPU c5 legs; scipy.sparse products are '
                            'single-threaded within each worker')
        cpu['c1_to_completion'] = cpu_c1(pg)
        log('cpu', json.dumps(cpu))
    agree = {n: agreement(outs32, v['outs']) for n, v in fp64.items()}
    el64 = fp64.get('fp64', {}).get('seconds')
    el64b = fp64.get('fp64_blocked', {}).get('seconds')
    out = dict(metric=METRIC, value=el32, unit='s', n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=el32 * 1e3, step_wall_s=wall32, higher_is_better=False, scaling='weak', vs_baseline=None,
               dtype='fp32 vectors / fp64 recurrences', data='synthetic',
               config=dict(workload=f'{a.config}: Chung-Lu n={inp["n"]}, m={inp["m"]}, sigma={inp["sigma"]}, '
                                    f's={list(inp["s_all"])}, q={inp["q"]}, p={inp["p"]}, N={inp["N"]}, '
                                    f'k={inp["k"]}, B={inp["B"]}, tol={inp["tol"]}',
                           headline_workload='c5 (see c5_greedy): Chung-Lu n=4M, m=34.7M, k=50, q=128, 32 poles, '
                                             'N=1e7, B=16',
                           precision='fp32-mixed', column_iterations=colits,
                           hipgraph=GRAPH, residual_replacement=replaced,
                           residual_replacement_note='FP64 residual replacement every 24 iterations (three '
                                                     '8-iteration windows of the blocked update; the c2 '
                                                     'configuration names 25)',
                           shared_bank=bool(outs32[0].get('shared_bank')),
                           s_schedule={'hybrid': f'{len(inp["s_all"])} s values: bank as one shared-seed solve (FP32-mixed; '
                                                 'row segments for FP64), re-rank as column groups',
                                       'columns': f'{len(inp["s_all"])} s values as column groups of one operator',
                                       'segments': f'{len(inp["s_all"])} s values as segments of one batched solve',
                                       'per-s': 'one greedy run per s value'}[a.schedule],
                           rhs=int(sum(o['rhs'] for o in outs32))),
               replacement_cost=repl_cost,
               fp64_seconds=el64, fp64_blocked_seconds=el64b,
               fp32_vs_fp64_speedup=(el64 / el32) if el64 else None,
               fp32_vs_fp64_blocked_speedup=(el64b / el32) if el64b else None,
               fp32_vs_fp64=agree.get('fp64'), fp32_vs_fp64_blocked=agree.get('fp64_blocked'),
               c5_greedy=c5, spmm_n4m=spmm, spmm_lattice_n4m=lattice, roofline=roofline, cpu_baseline=cpu,
               kernel_profile=prof, kernel_profile_note='one extra untimed FP32 step with HIP events around '
               'every kernel class (the timed steps carry events around k_opmul on one iteration in 8)')
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
