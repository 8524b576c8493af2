"""Batched form: POLAR greedy on many independent graphs at once.

The reference runs one graph at a time (realworld/run_real.py:114-133 loops over graphs, each
calling polar_gpu.polar, polar_gpu.py:180-254).  Small graphs leave the GPU idle: one CG iteration
of a 10k-row graph is a handful of microsecond-scale kernels.  Here all graphs become one
block-diagonal operator (`UnionGraph`), each graph a *segment* with its own poles and its own CG
scalars for every column; `polar_msolve_batched` advances every segment's solve with one launch per
kernel, skips finished segments and re-plans the blocks over the segments still iterating, and
(optionally) replays the CG iterations as hipGraphs.  The greedy rounds run in lockstep: one
batched bank solve, then per round one batched re-rank solve over all graphs' shortlists; the
per-graph proxy, shortlist, pick and Sherman-Morrison maintenance kernels work on views of the
batched caches.

Segment g's arithmetic is the single-graph arithmetic (same operator rows, same reductions), so
in FP64 every graph's selections, marginals, trace estimates and counters equal those of
`polar_gpu.polar` on that graph alone.
"""
import math
import time

import numpy as np
import torch

from . import _lib as L
from . import polar_gpu as pg


class UnionGraph(pg.Graph):
    """Block-diagonal union of graphs: graph g on rows [row0[g], row0[g+1])."""

    def __init__(self, graphs):
        graphs = list(graphs)
        ns = np.array([G.n for G in graphs], dtype=np.int64)
        row0 = np.zeros(len(graphs) + 1, dtype=np.int64)
        row0[1:] = np.cumsum(ns)
        edges = np.concatenate([G.edges + row0[g] for g, G in enumerate(graphs)])
        weights = np.concatenate([G.weights for G in graphs])
        super().__init__(int(row0[-1]), edges, 0.0, weights=weights)
        self.graphs = graphs
        self.row0 = row0
        # diag = deg + sigma of each row's own graph (the same doubles as each graph alone)
        self.sigma = np.repeat(np.array([G.sigma for G in graphs], dtype=np.float64), ns)
        self.dmax = max(G.dmax for G in graphs)

    def Lam(self, k):
        raise ValueError('a union graph has one Lam per segment: use graphs[g].Lam(k)')

    def segments(self, a0):
        """(rows, extra, longs, active) offsets of the segments, given the active-row offsets."""
        sc = self.schedule()
        lr = sc['long_rows']
        longs = np.searchsorted(lr, self.row0).astype(np.int64)
        base = np.r_[sc['long_base'].astype(np.int64), sc['n_partial']]
        extra = base[longs]
        return self.row0.astype(np.int64), extra, longs, np.asarray(a0, dtype=np.int64)


class Segments:
    """ctypes polar_segs over host offset arrays (kept alive by this object)."""

    def __init__(self, rows, extra, longs, active):
        self.arrays = [np.ascontiguousarray(a, dtype=np.int64) for a in (rows, extra, longs, active)]
        self.S = len(rows) - 1
        s = L.PolarSegs()
        s.S = self.S
        s.rows, s.extra, s.longs, s.active = (a.ctypes.data_as(L.vp) for a in self.arrays)
        self.c = s

    @property
    def a0(self):
        return self.arrays[3]


def msolve_batched(op, segs, W, tau, active, tol, maxit=100000, track=False, trace=False, out=None,
                   width=None, graph=None, freeze=True, block=None, replace=False, col_group=None, history=0):
    """Batched block multi-shift CG: segment g solves (M_g + tau[g, j]) U = W[rows of g, :R].

    op: Op of a UnionGraph; segs: Segments; W: (n, R) device; tau: (S, p) host or device; active:
    device int32 union row ids (segment g's at segs.a0[g]:segs.a0[g+1]).  Returns X (p, nA, R),
    h (p, S, R) or None, iters (S, R) host int64, and te (p, S, R) with trace=True.

    Column groups (one segment): col_group (R,) assigns column c to group col_group[c], which
    solves with the poles tau[col_group[c]] (tau is (G, p)) and the op's added edges tagged with
    that group; h, te are then (p, 1, R) and iters (1, R).

    history=T > 0 (one segment): POLAR_HISTORY with T iterations at most, as polar_gpu.msolve: X is
    None and a polar_gpu.History is appended to the results."""
    lib = L.require_gpu()
    dev = pg._device()
    dtype = op.dtype
    W = torch.as_tensor(W, device=dev).to(dtype)
    if W.stride(1) != 1:
        W = W.contiguous()
    n, R = W.shape
    S = segs.S
    if col_group is not None:
        assert S == 1
        col_group = np.ascontiguousarray(col_group, dtype=np.int32)
        assert col_group.shape == (R,)
        tau_d = torch.as_tensor(tau, dtype=torch.float64, device=dev)
        tau_d = tau_d.reshape(-1, tau_d.shape[-1]).contiguous()
    else:
        tau_d = torch.as_tensor(tau, dtype=torch.float64, device=dev).reshape(S, -1).contiguous()
    p = tau_d.shape[1]
    nA = int(segs.a0[-1])
    if history:
        assert S == 1 and out is None
        X = None
    else:
        X = out if out is not None else torch.empty((p, nA, R), dtype=dtype, device=dev)
        assert X.shape == (p, nA, R) and X.stride(2) == 1
    h = torch.empty((p, S, R), dtype=torch.float64, device=dev) if track else None
    te = torch.empty((p, S, R), dtype=torch.float64, device=dev) if trace else None
    iters = torch.empty((S, R), dtype=torch.int64, device=dev)
    if graph is None:
        graph = getattr(op, 'use_graph', False)
    flags = ((L.POLAR_TRACK if track else 0) | (L.POLAR_TRACE if trace else 0) |
             (0 if freeze else L.POLAR_NOFREEZE) | (L.POLAR_GRAPH if graph else 0) |
             (L.POLAR_BLOCK8 if pg._blocked(block, dtype) else 0) | (L.POLAR_REPLACE if replace else 0))
    s = op.c_op(dtype)
    step = min(R, pg._max_cols(dtype), width or R)
    if history:
        assert R <= pg._max_cols(dtype)
        step, maxit = R, int(history)
        flags |= L.POLAR_HISTORY
        ws_bytes = lib.polar_msolve_history_ws_bytes(s, R, p, nA, flags, maxit)
    else:
        ws_bytes = lib.polar_msolve_batched_ws_bytes(s, segs.c, step, p, flags)
    L.check(0 if ws_bytes > 0 else L.POLAR_E_ARG, 'polar_msolve_batched_ws_bytes')
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    act = active if nA else torch.zeros(1, dtype=torch.int32, device=dev)
    for lo in range(0, R, step):
        r = min(step, R - lo)
        hb = (h if r == R else torch.empty((p, S, r), dtype=torch.float64, device=dev)) if track else None
        tb = (te if r == R else torch.empty((p, S, r), dtype=torch.float64, device=dev)) if trace else None
        ib = iters if r == R else torch.empty((S, r), dtype=torch.int64, device=dev)
        if col_group is not None:
            segs.c.G = int(tau_d.shape[0])
            segs.c.col_group = col_group[lo:lo + r].ctypes.data_as(L.vp)
        rc = lib.polar_msolve_batched(s, segs.c, L.ptr(W[:, lo:]), W.stride(0), r, L.ptr(tau_d), p, L.ptr(act),
                                      float(tol), int(maxit), flags, None if X is None else L.ptr(X[:, :, lo:]),
                                      0 if X is None else X.stride(0), 0 if X is None else X.stride(1),
                                      L.ptr(hb), L.ptr(tb), L.ptr(ib), L.ptr(ws), ws_bytes, L.stream_ptr())
        if col_group is not None:
            segs.c.G, segs.c.col_group = 0, None
        L.check(rc, 'polar_msolve_batched')
        if r != R:
            iters[:, lo:lo + r] = ib
            if track:
                h[:, :, lo:lo + r] = hb
            if trace:
                te[:, :, lo:lo + r] = tb
    res = (X, h, iters.cpu().numpy())
    if trace:
        res = res + (te,)
    if history:
        res = res + (pg.History(op, s, R, p, nA, flags, maxit, ws, dtype),)
    return res


def _seq(x, S):
    if isinstance(x, (list, tuple, np.ndarray)):
        assert len(x) == S
        return list(x)
    return [x] * S


def polar_batched(graphs, cands, k, s, q=32, B=4, eps=1e-2, tol=1e-7, obj='frac', sketch=None, seed=2026,
                  kind='rademacher', poles=None, precision='fp32', graph=None, width=None, progress=None,
                  tune=None, block=None, union=None, replace=None):
    """POLAR greedy (polar_gpu.py:180-254, cache backend) on S graphs in lockstep.

    graphs: list of Graph; cands: list of (N_g, 2) candidate arrays; s: one value or one per graph;
    sketch: list of (Z, F, Gc) per graph (default: Rademacher with seed + g); poles: list of
    (tau, om) per graph (all with the same pole count) or None for `poles(obj, s_g, ...)`.
    graph: hipGraph replay (None: the tuned choice, else True); tune: a launch-shape dict from
    polar_amd.autotune or True to tune the batched operator here (before the timed region);
    union: a UnionGraph of these graphs built earlier (reused: its device arrays stay resident).
    replace: FP64 residual replacement in every solve; None = on for FP32, off for FP64.
    The same graph may appear several times (e.g. one graph under several s values).
    Returns one result dict per graph with the keys of polar_gpu.polar (`seconds` is the whole
    batch's wall time)."""
    lib = L.require_gpu()
    dev = pg._device()
    dtype = pg._dt(precision)
    replace = pg._replacing(replace, block, dtype)
    dcode = L.dtype_code(dtype)
    stream = L.stream_ptr()
    S = len(graphs)
    s_all = _seq(s, S)
    U = union if union is not None else UnionGraph(graphs)
    assert len(U.graphs) == S and all(a is b for a, b in zip(U.graphs, graphs))
    row0 = U.row0
    objc = 2 if obj == 'logdet' else 0
    want_trace = obj != 'logdet'
    taus, oms = [], []
    for g, G in enumerate(graphs):
        if poles is None:
            tau, om, _ = pg.poles(obj, s_all[g], G.sigma, G.Lam(k), eps)
        else:
            tau, om = (np.asarray(a, float) for a in poles[g])
        taus.append(tau)
        oms.append(om)
    p = len(taus[0])
    if any(len(t) != p for t in taus):
        raise ValueError('the batched form needs the same pole count for every graph')
    tau_all = torch.from_numpy(np.stack(taus)).to(dev)
    om_d = [torch.from_numpy(om).to(dev) for om in oms]
    cands = [np.asarray(c) for c in cands]
    if sketch is None:
        sketch = [pg.sketches(G, cands[g], q, seed + g, kind=kind, on_device='lazy' if kind == 'rademacher' else False,
                              precision=dtype) for g, G in enumerate(graphs)]
    q = sketch[0].q if isinstance(sketch[0], pg.DeviceSketch) else sketch[0][0].shape[1]
    n = U.n
    RHS = torch.empty((n, 2 * q), dtype=dtype, device=dev)
    actives, es, Gcs = [], [], []
    for g in range(S):
        Zg, Fg = RHS[row0[g]:row0[g + 1], :q], RHS[row0[g]:row0[g + 1], q:]
        if isinstance(sketch[g], pg.DeviceSketch):
            sketch[g].fill(Zg, Fg, pg.Op(graphs[g], precision=dtype))
            Gc = sketch[g].rows
        else:
            Z, F, Gc = sketch[g]
            Zg.copy_(torch.as_tensor(Z, device=dev).to(dtype))
            Fg.copy_(torch.as_tensor(F, device=dev).to(dtype))
        if isinstance(Gc, torch.Tensor):
            Gcs.append(Gc.to(dev, torch.float64))
        elif isinstance(Gc, pg.RademacherRows):
            Gcs.append(Gc)
        else:
            Gcs.append(np.asarray(Gc, dtype=np.float64))
        act = np.unique(cands[g])
        actives.append(act)
        es.append(np.searchsorted(act, cands[g]))
    a0 = np.zeros(S + 1, dtype=np.int64)
    a0[1:] = np.cumsum([len(a) for a in actives])
    segs = Segments(*U.segments(a0))
    active = torch.from_numpy(np.concatenate([a + row0[g] for g, a in enumerate(actives)]).astype(np.int32)).to(dev)
    nA = int(a0[-1])
    Ns = [len(c) for c in cands]
    ea = [torch.from_numpy(e[:, 0].astype(np.int32)).to(dev) for e in es]
    eb = [torch.from_numpy(e[:, 1].astype(np.int32)).to(dev) for e in es]
    op = pg.Op(U, precision=dtype)
    if tune is True:
        from . import autotune
        tune = autotune.tune(op, min(2 * q, pg._max_cols(dtype)), tau_all, active, segs=segs)
    if tune:
        from . import autotune
        autotune.apply(op, tune)
    if graph is None:
        graph = op.use_graph if tune else True
    spmv = np.zeros(S, dtype=np.int64)
    rhs = np.zeros(S, dtype=np.int64)
    live = [np.ones(N, bool) for N in Ns]
    dead = [torch.zeros(N, dtype=torch.uint8, device=dev) for N in Ns]
    score = [torch.empty(N, dtype=torch.float64, device=dev) for N in Ns]
    sel_ws = torch.empty(lib.polar_select_ws_bytes(max(Ns)), dtype=torch.uint8, device=dev)
    sel_out = torch.empty((S, 64), dtype=torch.int32, device=dev)
    coef = torch.empty((p, 2 * q + 3), dtype=torch.float64, device=dev)
    wX = torch.empty((p, max(len(a) for a in actives)), dtype=torch.float64, device=dev)
    wY = torch.empty_like(wX)
    wU = torch.empty_like(wX)
    gvec = torch.empty((S, max(B, 1) + 2), dtype=torch.float64, device=dev)
    # candidate sketch rows of the shortlists of the graphs whose sketch is the counter-hash kind,
    # generated on the GPU
    gc_dev = torch.empty((S, max(B, 1), q), dtype=torch.float64, device=dev) \
        if any(isinstance(Gc, pg.RademacherRows) for Gc in Gcs) else None
    gc_row = torch.empty(q, dtype=torch.float64, device=dev)
    seqs = [[] for _ in range(S)]
    margs = [[] for _ in range(S)]
    traces = [[] for _ in range(S)]
    tr_hist = [[] for _ in range(S)]
    pg._sync()
    t0 = time.perf_counter()

    XY = torch.empty((p, nA, 2 * q), dtype=dtype, device=dev)
    if width is None:
        width = pg.fit_width(n, p, nA, 2 * q, dtype, block=block, replace=replace)
    terms = [[] for _ in range(S)]
    for lo in range(0, 2 * q, width):
        hi = min(2 * q, lo + width)
        res = msolve_batched(op, segs, RHS[:, lo:hi], tau_all, active, tol, trace=want_trace and lo < q,
                             out=XY[:, :, lo:hi], graph=graph, block=block, replace=replace)
        spmv += res[2].sum(axis=1)
        rhs += hi - lo
        if want_trace and lo < q:
            ncol = min(q, hi) - lo
            te = res[3][:, :, :ncol].cpu().numpy()
            for g in range(S):
                terms[g].append((oms[g][:, None] * te[:, g, :]).ravel())
    X, Y = XY[:, :, :q], XY[:, :, q:]
    hs, ts = [], []
    for g in range(S):
        Xg, Yg = X[:, a0[g]:], Y[:, a0[g]:]
        h = torch.empty((p, Ns[g]), dtype=torch.float64, device=dev)
        t = torch.empty((p, Ns[g]), dtype=torch.float64, device=dev)
        L.check(lib.polar_pair_sqnorm(dcode, L.ptr(Xg), Xg.stride(0), Xg.stride(1), p, q, L.ptr(ea[g]),
                                      L.ptr(eb[g]), Ns[g], L.ptr(h), stream), 'pair_sqnorm')
        L.check(lib.polar_pair_sqnorm(dcode, L.ptr(Yg), Yg.stride(0), Yg.stride(1), p, q, L.ptr(ea[g]),
                                      L.ptr(eb[g]), Ns[g], L.ptr(t), stream), 'pair_sqnorm')
        hs.append(h)
        ts.append(t)
        traces[g].append(math.fsum(np.concatenate(terms[g])) if want_trace else None)
    if progress is not None:
        progress('bank', dict(seconds=time.perf_counter() - t0, spmv=int(spmv.sum())))

    for rnd in range(k):
        # every graph's proxy and top-B on the device, then one read-back of all shortlists
        Brs = [min(B, int(live[g].sum())) for g in range(S)]
        for g in range(S):
            if Brs[g] == 0:                   # candidates exhausted: this graph stops selecting
                continue
            L.check(lib.polar_proxy(objc, L.ptr(hs[g]), L.ptr(ts[g]), L.ptr(tau_all[g]), L.ptr(om_d[g]), p, Ns[g],
                                    graphs[g].sigma, L.ptr(dead[g]), L.ptr(score[g]), stream), 'proxy')
            L.check(lib.polar_select_topb(L.ptr(score[g]), Ns[g], Brs[g], L.ptr(sel_out[g]), None,
                                          L.ptr(sel_ws), sel_ws.numel(), stream), 'select_topb')
            if isinstance(Gcs[g], pg.RademacherRows):
                Gcs[g].fill_device(sel_out[g, :Brs[g]].to(torch.int64), gc_dev[g])
        Bm = max(Brs)
        if Bm == 0:
            break
        sel_h = sel_out.cpu().numpy()
        shorts = [np.ascontiguousarray(sel_h[g, :Brs[g]]) for g in range(S)]
        # re-rank right-hand sides (union rows, column = position in the graph's shortlist) and the
        # shortlists' endpoint slots, uploaded together
        up = []
        for g in range(S):
            c = cands[g][shorts[g]]
            up.append(np.stack([c[:, 0] + row0[g], c[:, 1] + row0[g], np.arange(Brs[g]),
                                es[g][shorts[g], 0], es[g][shorts[g], 1]]))
        idx = torch.from_numpy(np.concatenate(up, axis=1).astype(np.int64)).to(dev)
        ia, ib, ic = idx[0], idx[1], idx[2]
        sab = idx[3:].to(torch.int32)
        off = np.r_[0, np.cumsum(Brs)]
        Wb = torch.zeros((n, Bm), dtype=dtype, device=dev)
        Wb.index_put_((ia, ic), torch.ones(len(ic), dtype=dtype, device=dev))
        Wb.index_put_((ib, ic), torch.full((len(ic),), -1.0, dtype=dtype, device=dev))
        Ub, hh, it = msolve_batched(op, segs, Wb, tau_all, active, tol, track=obj != 'logdet', graph=graph,
                                    block=block, replace=replace)[:3]
        spmv += it.sum(axis=1)
        # every graph's re-rank pick on the device, then one read-back of all winners
        for g in range(S):
            if Brs[g]:
                Ug = Ub[:, a0[g]:]
                hg = hh[:, g, :Brs[g]].contiguous() if hh is not None else None
                L.check(lib.polar_rerank_pick(objc, dcode, L.ptr(Ug), Ug.stride(0), Ug.stride(1), p, Brs[g],
                                              L.ptr(sab[0, off[g]:]), L.ptr(sab[1, off[g]:]), L.ptr(hg),
                                              L.ptr(om_d[g]), L.ptr(gvec[g]), None, None, stream), 'rerank_pick')
        gv_h = gvec.cpu().numpy()
        for g in range(S):
            Br = Brs[g]
            if Br == 0:
                continue
            rhs[g] += Br
            short = shorts[g]
            Ug = Ub[:, a0[g]:]
            j = 0 if Br == 1 else int(gv_h[g, Br])
            e = int(short[j])
            u = Ug[:, :, j]
            margs[g].append(float(gv_h[g, Br + 1]))
            seqs[g].append(e)
            live[g][e] = False
            dead[g][e] = 1
            aa, bb = (int(x) for x in es[g][e])
            a, b = (int(x) for x in cands[g][e])
            Gc = Gcs[g]
            if isinstance(Gc, pg.RademacherRows):
                gcr = gc_dev[g, j]
            else:
                gcr = gc_row
                if isinstance(Gc, torch.Tensor):
                    gc_row.copy_(Gc[e])
                else:
                    gc_row.copy_(torch.from_numpy(np.ascontiguousarray(Gc[e], dtype=np.float64)))
            if rnd + 1 < k:
                Xg, Yg = X[:, a0[g]:], Y[:, a0[g]:]
                nAg = int(a0[g + 1] - a0[g])
                L.check(lib.polar_maint_prep(dcode, L.ptr(Xg), L.ptr(Yg), Xg.stride(0), Xg.stride(1), p, q, L.ptr(u),
                                             u.stride(0), u.stride(1), aa, bb, L.ptr(gcr), L.ptr(coef), stream),
                        'maint_prep')
                L.check(lib.polar_maint_rank1(dcode, L.ptr(Xg), L.ptr(Yg), Xg.stride(0), Xg.stride(1), p, nAg, q,
                                              L.ptr(u), u.stride(0), u.stride(1), L.ptr(coef), L.ptr(wX), L.ptr(wY),
                                              L.ptr(wU), stream), 'maint_rank1')
                L.check(lib.polar_maint_cache(p, Ns[g], L.ptr(ea[g]), L.ptr(eb[g]), L.ptr(coef), q, L.ptr(wX),
                                              L.ptr(wY), L.ptr(wU), L.ptr(hs[g]), L.ptr(ts[g]), stream),
                        'maint_cache')
                if want_trace:
                    tr_hist[g].append(coef[:, 2 * q:2 * q + 2].clone())
            Fv = RHS[:, q:]
            L.check(lib.polar_rows_axpy(dcode, L.ptr(Fv), Fv.stride(0), q, a + int(row0[g]), b + int(row0[g]),
                                        L.ptr(gcr), stream), 'rows_axpy')
            op.add(a + int(row0[g]), b + int(row0[g]))
        if progress is not None:
            progress('round', dict(round=rnd, edges=[sq[-1] if sq else None for sq in seqs], seconds=time.perf_counter() - t0,
                                   spmv=int(spmv.sum())))
    pg._sync()
    seconds = time.perf_counter() - t0
    out = []
    for g in range(S):
        tr = traces[g]
        if want_trace:
            for dn in tr_hist[g]:
                dn_h = dn.cpu().numpy()
                tr.append(tr[-1] - math.fsum(oms[g] * dn_h[:, 1] / dn_h[:, 0]))
        out.append(dict(sequence=seqs[g], seconds=seconds, spmv=int(spmv[g]), rhs=int(rhs[g]), p=p,
                        rounds=len(seqs[g]), truncated=len(seqs[g]) < k, marginals=margs[g],
                        trace_estimate=tr[:max(len(seqs[g]), 1)], replace=replace,
                        precision='fp64' if dtype == torch.float64 else 'fp32'))
    return out


def single_segment(G, nA):
    """Segments of a plain graph (one segment) with nA active rows."""
    sc = G.schedule()
    return Segments(np.array([0, G.n]), np.array([0, len(sc['extra'])]), np.array([0, len(sc['long_rows'])]),
                    np.array([0, nA]))


def polar_columns(G, cand, k, s_list, q=64, B=8, eps=1e-2, tol=1e-7, obj='frac', sketch=None, seed=2026,
                  kind='rademacher', poles=None, precision='fp32', graph=None, width=None, block=None,
                  replace=None, progress=None, bank_union=None, shared_bank=None, history=None):
    """POLAR greedy (polar_gpu.py:180-254) for several s values on ONE graph, in lockstep, as column
    groups of one operator: group g is the run for s_list[g], with its own poles, its own bank
    columns [Z | F_g], its own caches and its own selected edges (patch entries tagged g).  Every
    solve of a phase is one launch sequence that reads the adjacency once for all groups.  Results
    equal polar_gpu.polar run once per s (same sketch): one dict per s value.

    bank_union: a UnionGraph of len(s_list) copies of G; the bank then runs as one segmented solve
    on it (the groups' 2q columns side by side in rows, finished groups leave the grid) and the
    re-rank solves as column groups (same results either way).  replace: FP64 residual replacement
    in every solve; None = on for FP32, off for FP64.

    history: the re-rank solve with the blocked update forms the shifted solutions only on the
    shortlists' endpoint rows and the winners' columns (polar_gpu.polar's `history`); None = when it
    fits, False = never.

    shared_bank: the groups' bank right-hand sides [Z | F] are the same, so one seed CG can serve all
    of them: the seed is the smallest pole of all groups and every group's poles are its shifted
    systems (multi-shift CG), one pass over the adjacency and one set of vector streams per
    iteration for all s values.  Each pole's solution still meets the tolerance, but the Krylov
    seed differs from the per-s run's, so results agree with polar_gpu.polar within the CG
    tolerance rather than bit for bit.  None = on for FP32 (whose results carry FP32 rounding
    anyway), off for FP64 (bit-identical to the per-s runs and to the oracle).  Per-group `spmv`
    then counts the shared bank solve's column-iterations in every group."""
    lib = L.require_gpu()
    dev = pg._device()
    dtype = pg._dt(precision)
    replace = pg._replacing(replace, block, dtype)
    if shared_bank is None:
        shared_bank = dtype == torch.float32
    dcode = L.dtype_code(dtype)
    stream = L.stream_ptr()
    Gn = len(s_list)
    objc = 2 if obj == 'logdet' else 0
    want_trace = obj != 'logdet'
    if poles is None:
        poles = [pg.poles(obj, s, G.sigma, G.Lam(k), eps)[:2] for s in s_list]
    taus = [np.asarray(t, float) for t, _ in poles]
    oms = [np.asarray(o, float) for _, o in poles]
    p = len(taus[0])
    if any(len(t) != p for t in taus):
        raise ValueError('column groups need the same pole count for every s')
    tau_all = torch.from_numpy(np.stack(taus)).to(dev)
    om_d = [torch.from_numpy(o).to(dev) for o in oms]
    cand = np.asarray(cand)
    N = len(cand)
    if sketch is None:
        sketch = pg.sketches(G, cand, q, seed, kind=kind, on_device='lazy' if kind == 'rademacher' else False,
                             precision=dtype)
    n = G.n
    op = pg.Op(G, precision=dtype)
    W2 = 2 * q if not isinstance(sketch, pg.DeviceSketch) else 2 * sketch.q
    q = W2 // 2
    RHS = torch.empty((n, W2 * Gn), dtype=dtype, device=dev)        # [Z | F_g] per group
    if isinstance(sketch, pg.DeviceSketch):
        sketch.fill(RHS[:, :q], RHS[:, q:W2], op)
        Gc = sketch.rows
    else:
        Z, F, Gc = sketch
        RHS[:, :q].copy_(torch.as_tensor(Z, device=dev).to(dtype))
        RHS[:, q:W2].copy_(torch.as_tensor(F, device=dev).to(dtype))
    for g in range(1, Gn):
        RHS[:, g * W2:(g + 1) * W2].copy_(RHS[:, :W2])
    if isinstance(Gc, torch.Tensor):
        Gc = Gc.to(dev, torch.float64)
    elif not isinstance(Gc, pg.RademacherRows):
        Gc = np.asarray(Gc, dtype=np.float64)
    active_h = np.unique(cand)
    es_h = np.searchsorted(active_h, cand)
    nA = len(active_h)
    active = torch.from_numpy(active_h.astype(np.int32)).to(dev)
    ea = torch.from_numpy(es_h[:, 0].astype(np.int32)).to(dev)
    eb = torch.from_numpy(es_h[:, 1].astype(np.int32)).to(dev)
    segs = single_segment(G, nA)
    spmv = np.zeros(Gn, dtype=np.int64)
    rhs = np.zeros(Gn, dtype=np.int64)
    live = [np.ones(N, bool) for _ in range(Gn)]
    dead = [torch.zeros(N, dtype=torch.uint8, device=dev) for _ in range(Gn)]
    score = torch.empty(N, dtype=torch.float64, device=dev)
    sel_ws = torch.empty(lib.polar_select_ws_bytes(N), dtype=torch.uint8, device=dev)
    sel_out = torch.empty((Gn, 64), dtype=torch.int32, device=dev)
    coef = torch.empty((p, 2 * q + 3), dtype=torch.float64, device=dev)
    wX = torch.empty((p, nA), dtype=torch.float64, device=dev)
    wY = torch.empty_like(wX)
    wU = torch.empty_like(wX)
    gvec = torch.empty((Gn, max(B, 1) + 2), dtype=torch.float64, device=dev)
    # candidate sketch rows of the shortlists, generated on the GPU (counter-hash kind)
    gc_dev = torch.empty((Gn, max(B, 1), q), dtype=torch.float64, device=dev) \
        if isinstance(Gc, pg.RademacherRows) else None
    gc_row = torch.empty(q, dtype=torch.float64, device=dev)
    # endpoint row indices 0 / 1 of a (p, 2, R) block of endpoint values (history re-rank)
    z01 = torch.arange(2, dtype=torch.int32, device=dev)[:, None].repeat(1, max(B, 1)).contiguous()
    seqs = [[] for _ in range(Gn)]
    margs = [[] for _ in range(Gn)]
    traces = [[] for _ in range(Gn)]
    tr_hist = [[] for _ in range(Gn)]
    pg._sync()
    t0 = time.perf_counter()

    terms = [[] for _ in range(Gn)]
    if shared_bank and Gn > 1:
        # group-major pole layout, the group holding the smallest pole first (tau[0] is the seed)
        order = sorted(range(Gn), key=lambda g: (taus[g][0], g))[:1]
        order += [g for g in range(Gn) if g != order[0]]
        tau_cat = np.concatenate([taus[g] for g in order])
        pc = Gn * p
        XYs = torch.empty((pc, nA, W2), dtype=dtype, device=dev)
        ws = width or pg.fit_width(n, pc, nA, W2, dtype, block=block, replace=replace)
        for lo in range(0, W2, ws):
            hi = min(W2, lo + ws)
            res = pg.msolve(op, RHS[:, lo:hi], tau_cat, active, tol, trace=want_trace and lo < q,
                            out=XYs[:, :, lo:hi], graph=graph, block=block, replace=replace)
            cols = int(res[2].sum())
            for g in range(Gn):
                spmv[g] += cols
                rhs[g] += hi - lo
            if want_trace and lo < q:
                te = res[3][:, :min(q, hi) - lo].cpu().numpy()
                for slot, g in enumerate(order):
                    terms[g].append((oms[g][:, None] * te[slot * p:(slot + 1) * p]).ravel())
        Xs, Ys = [None] * Gn, [None] * Gn
        for slot, g in enumerate(order):
            Xs[g] = XYs[slot * p:(slot + 1) * p, :, :q]
            Ys[g] = XYs[slot * p:(slot + 1) * p, :, q:]
    elif bank_union is not None:
        U = bank_union
        assert len(U.graphs) == Gn and all(x is G for x in U.graphs)
        a0 = np.arange(Gn + 1, dtype=np.int64) * nA
        segs_u = Segments(*U.segments(a0))
        act_u = torch.cat([active + int(U.row0[g]) for g in range(Gn)])
        op_u = pg.Op(U, precision=dtype)
        RHS_u = torch.empty((U.n, W2), dtype=dtype, device=dev)
        for g in range(Gn):
            RHS_u[U.row0[g]:U.row0[g + 1]].copy_(RHS[:, g * W2:(g + 1) * W2])
        XYu = torch.empty((p, Gn * nA, W2), dtype=dtype, device=dev)
        wu = width or pg.fit_width(U.n, p, Gn * nA, W2, dtype, block=block, replace=replace)
        for lo in range(0, W2, wu):
            hi = min(W2, lo + wu)
            res = msolve_batched(op_u, segs_u, RHS_u[:, lo:hi], tau_all, act_u, tol, trace=want_trace and lo < q,
                                 out=XYu[:, :, lo:hi], graph=graph, block=block, replace=replace)
            for g in range(Gn):
                spmv[g] += int(res[2][g].sum())
                rhs[g] += hi - lo
                if want_trace and lo < q:
                    te = res[3][:, g, :min(q, hi) - lo].cpu().numpy()
                    terms[g].append((oms[g][:, None] * te).ravel())
        del RHS_u
        Xs = [XYu[:, g * nA:(g + 1) * nA, :q] for g in range(Gn)]
        Ys = [XYu[:, g * nA:(g + 1) * nA, q:] for g in range(Gn)]
    else:
        C = W2 * Gn
        XY = torch.empty((p, nA, C), dtype=dtype, device=dev)
        if width is None:
            width = pg.fit_width(n, p, nA, C, dtype, block=block, replace=replace)
        col_of = np.arange(C) // W2                                  # group of every bank column
        for lo in range(0, C, width):
            hi = min(C, lo + width)
            res = msolve_batched(op, segs, RHS[:, lo:hi], tau_all, active, tol, trace=want_trace,
                                 out=XY[:, :, lo:hi], graph=graph, block=block, replace=replace,
                                 col_group=col_of[lo:hi])
            it = res[2][0]
            te = res[3][:, 0, :].cpu().numpy() if want_trace else None
            for c in range(lo, hi):
                g = col_of[c]
                spmv[g] += it[c - lo]
                rhs[g] += 1
                if want_trace and c - g * W2 < q:
                    terms[g].append(oms[g] * te[:, c - lo])
        Xs = [XY[:, :, g * W2:g * W2 + q] for g in range(Gn)]
        Ys = [XY[:, :, g * W2 + q:(g + 1) * W2] for g in range(Gn)]
    bank_spmv = spmv.copy()
    hs, ts = [], []
    for g in range(Gn):
        h = torch.empty((p, N), dtype=torch.float64, device=dev)
        t = torch.empty((p, N), dtype=torch.float64, device=dev)
        L.check(lib.polar_pair_sqnorm(dcode, L.ptr(Xs[g]), Xs[g].stride(0), Xs[g].stride(1), p, q, L.ptr(ea), L.ptr(eb),
                                      N, L.ptr(h), stream), 'pair_sqnorm')
        L.check(lib.polar_pair_sqnorm(dcode, L.ptr(Ys[g]), Ys[g].stride(0), Ys[g].stride(1), p, q, L.ptr(ea), L.ptr(eb),
                                      N, L.ptr(t), stream), 'pair_sqnorm')
        hs.append(h)
        ts.append(t)
        traces[g].append(math.fsum(np.concatenate(terms[g])) if want_trace else None)
    if progress is not None:
        progress('bank', dict(seconds=time.perf_counter() - t0, spmv=int(spmv.sum())))

    hist_ok = True                     # cleared once a history solve outruns its history
    for rnd in range(k):
        # every group's proxy and top-B on the device, then one read-back of all shortlists
        Brs = [min(B, int(live[g].sum())) for g in range(Gn)]
        for g in range(Gn):
            if Brs[g]:
                L.check(lib.polar_proxy(objc, L.ptr(hs[g]), L.ptr(ts[g]), L.ptr(tau_all[g]), L.ptr(om_d[g]), p, N,
                                        G.sigma, L.ptr(dead[g]), L.ptr(score), stream), 'proxy')
                L.check(lib.polar_select_topb(L.ptr(score), N, Brs[g], L.ptr(sel_out[g]), None,
                                              L.ptr(sel_ws), sel_ws.numel(), stream), 'select_topb')
        if gc_dev is not None and any(Brs):
            keys = sel_out[:, :max(B, 1)].to(torch.int64).reshape(-1)
            Gc.fill_device(keys, gc_dev.view(-1, q))
        sel_h = sel_out.cpu().numpy()
        shorts = [np.ascontiguousarray(sel_h[g, :Brs[g]]) for g in range(Gn)]
        off = np.r_[0, np.cumsum(Brs)]
        Rt = int(off[-1])
        if Rt == 0:
            break
        allsh = np.concatenate(shorts)
        # re-rank right-hand sides and the shortlists' endpoint slots, uploaded together
        idx = torch.from_numpy(np.stack([cand[allsh, 0], cand[allsh, 1], es_h[allsh, 0], es_h[allsh, 1]])).to(dev)
        ia, ib = idx[0], idx[1]
        sab = idx[2:].to(torch.int32)
        ic = torch.arange(Rt, device=dev)
        Wb = torch.zeros((n, Rt), dtype=dtype, device=dev)
        Wb.index_put_((ia, ic), torch.ones(Rt, dtype=dtype, device=dev))
        Wb.index_put_((ib, ic), torch.full((Rt,), -1.0, dtype=dtype, device=dev))
        cg = np.concatenate([np.full(len(sh), g, np.int32) for g, sh in enumerate(shorts)])
        H = None
        hcap = pg.history_cap(n, Rt, dtype, block) if history is not False and hist_ok else 0
        if hcap:
            # FP32: keep the residuals, form X_j only at the shortlisted endpoints and the winners
            try:
                _, hh, it, H = msolve_batched(op, segs, Wb, tau_all, active, tol, track=obj != 'logdet',
                                              graph=graph, block=block, replace=replace, col_group=cg, history=hcap)
            except RuntimeError as ex:
                if 'iteration cap' not in str(ex):
                    raise
                hist_ok = False                # the later re-rank solves need about as many
        if H is None:
            Ub, hh, it = msolve_batched(op, segs, Wb, tau_all, active, tol, track=obj != 'logdet', graph=graph,
                                        block=block, replace=replace, col_group=cg)[:3]
        else:
            Ub = H.rows(idx[:2])                       # (p, 2, Rt): U at the two endpoints of every column
        # every group's re-rank pick on the device, then one read-back of all winners
        for g in range(Gn):
            if Brs[g]:
                Ug = Ub[:, :, off[g]:]
                hg = hh[:, 0, off[g]:off[g + 1]].contiguous() if hh is not None else None
                sa, sb = (sab[0, off[g]:], sab[1, off[g]:]) if H is None else (z01[0], z01[1])
                L.check(lib.polar_rerank_pick(objc, dcode, L.ptr(Ug), Ug.stride(0), Ug.stride(1), p, Brs[g],
                                              L.ptr(sa), L.ptr(sb), L.ptr(hg), L.ptr(om_d[g]),
                                              L.ptr(gvec[g]), None, None, stream), 'rerank_pick')
        gv_h = gvec.cpu().numpy()
        wins = {g: (0 if Brs[g] == 1 else int(gv_h[g, Brs[g]])) for g in range(Gn) if Brs[g]}
        if H is not None and rnd + 1 < k and wins:
            # the winners' columns on every active row, for the maintenance
            wg = sorted(wins)
            Uw = H.cols(torch.from_numpy(np.array([off[g] + wins[g] for g in wg], dtype=np.int32)).to(dev), active)
            Uwin = {g: Uw[i] for i, g in enumerate(wg)}
        for g in range(Gn):
            Br = Brs[g]
            if Br == 0:
                continue
            spmv[g] += int(it[0, off[g]:off[g + 1]].sum())
            rhs[g] += Br
            short = shorts[g]
            j = wins[g]
            e = int(short[j])
            u = Ub[:, :, off[g] + j] if H is None else (Uwin[g] if rnd + 1 < k else None)
            margs[g].append(float(gv_h[g, Br + 1]))
            seqs[g].append(e)
            live[g][e] = False
            dead[g][e] = 1
            aa, bb = (int(x) for x in es_h[e])
            a, b = (int(x) for x in cand[e])
            if gc_dev is not None:
                gcr = gc_dev[g, j]
            else:
                gcr = gc_row
                if isinstance(Gc, torch.Tensor):
                    gc_row.copy_(Gc[e])
                else:
                    gc_row.copy_(torch.from_numpy(np.ascontiguousarray(Gc[e], dtype=np.float64)))
            if rnd + 1 < k:
                Xg, Yg = Xs[g], Ys[g]
                L.check(lib.polar_maint_prep(dcode, L.ptr(Xg), L.ptr(Yg), Xg.stride(0), Xg.stride(1), p, q, L.ptr(u),
                                             u.stride(0), u.stride(1), aa, bb, L.ptr(gcr), L.ptr(coef), stream),
                        'maint_prep')
                L.check(lib.polar_maint_rank1(dcode, L.ptr(Xg), L.ptr(Yg), Xg.stride(0), Xg.stride(1), p, nA, q,
                                              L.ptr(u), u.stride(0), u.stride(1), L.ptr(coef), L.ptr(wX), L.ptr(wY),
                                              L.ptr(wU), stream), 'maint_rank1')
                L.check(lib.polar_maint_cache(p, N, L.ptr(ea), L.ptr(eb), L.ptr(coef), q, L.ptr(wX), L.ptr(wY),
                                              L.ptr(wU), L.ptr(hs[g]), L.ptr(ts[g]), stream), 'maint_cache')
                if want_trace:
                    tr_hist[g].append(coef[:, 2 * q:2 * q + 2].clone())
            Fv = RHS[:, g * W2 + q:(g + 1) * W2]
            L.check(lib.polar_rows_axpy(dcode, L.ptr(Fv), Fv.stride(0), q, a, b, L.ptr(gcr), stream), 'rows_axpy')
            op.add(a, b, group=g)
        if progress is not None:
            progress('round', dict(round=rnd, edges=[sq[-1] if sq else None for sq in seqs],
                                   seconds=time.perf_counter() - t0, spmv=int(spmv.sum())))
    pg._sync()
    seconds = time.perf_counter() - t0
    out = []
    for g in range(Gn):
        tr = traces[g]
        if want_trace:
            for dn in tr_hist[g]:
                dn_h = dn.cpu().numpy()
                tr.append(tr[-1] - math.fsum(oms[g] * dn_h[:, 1] / dn_h[:, 0]))
        out.append(dict(sequence=seqs[g], seconds=seconds, spmv=int(spmv[g]), rhs=int(rhs[g]), p=p,
                        rounds=len(seqs[g]), truncated=len(seqs[g]) < k, marginals=margs[g],
                        trace_estimate=tr[:max(len(seqs[g]), 1)], replace=replace,
                        shared_bank=bool(shared_bank and Gn > 1), bank_spmv=int(bank_spmv[g]),
                        precision='fp64' if dtype == torch.float64 else 'fp32'))
    return out
