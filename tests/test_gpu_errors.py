"""Argument checks of the C-ABI and degenerate inputs of the HIP path.

Every entry point checks its sizes, pointers and dtype on the host and returns POLAR_E_ARG with a
message (raised as polar_amd._lib.PolarError) before it launches anything.  Degenerate but valid
inputs run and give the closed-form results: a graph without edges (every row is its diagonal, so
each shifted system is a multiple of the identity), zero-row calls, and the clear of the re-rank
right-hand side at sizes that take either path of the fill kernel."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from oracle import polar_oracle as O  # noqa: E402
from polar_amd import _lib as L  # noqa: E402
from polar_amd import graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _small_op(precision='fp64'):
    n = 64
    e = graphs.barabasi_albert(n, 2, 1)
    op = pg.Op(pg.Graph(n, e, 0.5), precision=precision)
    dt = torch.float64 if precision == 'fp64' else torch.float32
    return op, op.c_op(dt), dt, n


def _expect_arg(rc, match):
    assert rc == L.POLAR_E_ARG
    with pytest.raises(L.PolarError, match=match):
        L.check(rc, 'call')


def test_spmm_rejects_bad_widths():
    lib = L.require_gpu()
    op, s, dt, n = _small_op()
    X = torch.zeros((n, 300), dtype=dt, device='cuda')
    Y = torch.zeros_like(X)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device='cuda')
    sp = L.stream_ptr()
    _expect_arg(lib.polar_spmm(s, 0.0, L.ptr(X), 300, L.ptr(Y), 300, 0, L.ptr(ws), ws.numel(), sp), 'bad R/ld')
    _expect_arg(lib.polar_spmm(s, 0.0, L.ptr(X), 4, L.ptr(Y), 300, 8, L.ptr(ws), ws.numel(), sp), 'bad R/ld')
    # 300 FP64 columns need 150 lanes of 2 columns: more than one wave
    _expect_arg(lib.polar_spmm(s, 0.0, L.ptr(X), 300, L.ptr(Y), 300, 300, L.ptr(ws), ws.numel(), sp), 'too wide')


@pytest.mark.parametrize('field,value,match', [('dtype', 7, 'invalid polar_op'), ('n', 0, 'invalid polar_op'),
                                               ('grid_max', 70000, 'grid_max'), ('cpl_max', 3, 'cpl_max'),
                                               ('chunk', 4, 'chunk length')])
def test_operator_descriptor_checks(field, value, match):
    lib = L.require_gpu()
    op, s, dt, n = _small_op()
    bad = L.PolarOp.from_buffer_copy(s)
    setattr(bad, field, value)
    assert lib.polar_spmm_ws_bytes(ctypes.byref(bad), 8) == -1
    X = torch.zeros((n, 8), dtype=dt, device='cuda')
    Y = torch.zeros_like(X)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device='cuda')
    _expect_arg(lib.polar_spmm(ctypes.byref(bad), 0.0, L.ptr(X), 8, L.ptr(Y), 8, 8, L.ptr(ws), ws.numel(),
                               L.stream_ptr()), match)


def test_msolve_rejects_bad_arguments():
    lib = L.require_gpu()
    op, s, dt, n = _small_op()
    W = torch.zeros((n, 520), dtype=dt, device='cuda')
    tau = torch.zeros(2, dtype=torch.float64, device='cuda')
    act = torch.arange(n, dtype=torch.int32, device='cuda')
    X = torch.zeros(8, dtype=dt, device='cuda')
    its = torch.zeros(520, dtype=torch.int64, device='cuda')
    ws = torch.empty(256, dtype=torch.uint8, device='cuda')

    def call(R, p, nA):
        return lib.polar_msolve(s, L.ptr(W), 520, R, L.ptr(tau), p, L.ptr(act), nA, 1e-8, 100, 0, L.ptr(X), 0, 0,
                                None, None, L.ptr(its), L.ptr(ws), ws.numel(), L.stream_ptr())
    _expect_arg(call(513, 2, n), 'exceeds 512 columns')
    _expect_arg(call(0, 2, n), 'bad msolve arguments')
    _expect_arg(call(8, 0, n), 'bad msolve arguments')
    _expect_arg(call(8, 2, 0), 'bad msolve arguments')


def test_greedy_entry_checks():
    lib = L.require_gpu()
    sp = L.stream_ptr()
    d = torch.zeros(4096, dtype=torch.float64, device='cuda')
    i = torch.zeros(64, dtype=torch.int32, device='cuda')
    ws = torch.empty(lib.polar_select_ws_bytes(100), dtype=torch.uint8, device='cuda')
    P = L.ptr
    _expect_arg(lib.polar_pair_sqnorm(5, P(d), 8, 8, 1, 8, P(i), P(i), 4, P(d), sp), 'pair_sqnorm')
    _expect_arg(lib.polar_proxy(0, P(d), P(d), P(d), P(d), 2, -1, 0.5, None, P(d), sp), 'proxy')
    _expect_arg(lib.polar_select_topb(P(d), 100, 65, P(i), None, P(ws), ws.numel(), sp), 'select')
    _expect_arg(lib.polar_select_topb(P(d), 3, 4, P(i), None, P(ws), ws.numel(), sp), 'select')
    _expect_arg(lib.polar_select_topb(P(d), 100, 4, P(i), None, P(ws), 16, sp), 'workspace too small')
    _expect_arg(lib.polar_maint_rank1(0, P(d), P(d), 8, 8, 2, -1, 8, P(d), 8, 8, P(d), P(d), P(d), P(d), sp),
                'maint_rank1')
    _expect_arg(lib.polar_maint_cache(4096, 10, P(i), P(i), P(d), 8, P(d), P(d), P(d), P(d), P(d), sp),
                'exceed the LDS tile')
    _expect_arg(lib.polar_maint_cache(2, -5, P(i), P(i), P(d), 8, P(d), P(d), P(d), P(d), P(d), sp), 'maint_cache')
    _expect_arg(lib.polar_rows_axpy(0, P(d), 4, 8, 0, 1, P(d), sp), 'rows_axpy')
    _expect_arg(lib.polar_rademacher(3, 1, 0, None, 4, 8, 1.0, P(d), 8, sp), 'rademacher')
    _expect_arg(lib.polar_rademacher(0, 1, 0, None, 4, 8, 1.0, P(d), 4, sp), 'rademacher')
    # zero rows / zero pairs are valid no-ops
    assert lib.polar_maint_rank1(0, P(d), P(d), 8, 8, 2, 0, 8, P(d), 8, 8, P(d), P(d), P(d), P(d), sp) == 0
    assert lib.polar_maint_cache(2, 0, P(i), P(i), P(d), 8, P(d), P(d), P(d), P(d), P(d), sp) == 0
    assert lib.polar_proxy(0, P(d), P(d), P(d), P(d), 2, 0, 0.5, None, P(d), sp) == 0
    torch.cuda.synchronize()


@pytest.mark.parametrize('dt,n,Br', [(torch.float32, 5, 3), (torch.float64, 5, 3), (torch.float32, 1001, 7),
                                     (torch.float64, 1000, 8)])
def test_pair_rhs_clears_and_sets_every_size(dt, n, Br):
    """W starts as NaN; afterwards it is exactly e_a - e_b per column (the clear runs as 4-byte
    stores when n*Br*size is not a multiple of 16 and as 16-byte stores otherwise)."""
    lib = L.require_gpu()
    rng = np.random.default_rng(n + Br)
    a = rng.choice(n, Br, replace=False).astype(np.int32)
    b = np.array([(x + 1) % n for x in a], dtype=np.int32)
    W = torch.full((n, Br), float('nan'), dtype=dt, device='cuda')
    a_d, b_d = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    L.check(lib.polar_pair_rhs(L.dtype_code(dt), n, Br, L.ptr(a_d), L.ptr(b_d), L.ptr(W), L.stream_ptr()), 'pair_rhs')
    want = np.zeros((n, Br))
    want[a, np.arange(Br)] = 1.0
    want[b, np.arange(Br)] = -1.0
    assert np.array_equal(W.cpu().double().numpy(), want)


def test_edgeless_graph_operator_and_solve():
    """No edges: (M0 + shift) x = (sigma + shift) x row by row, and every shifted system is solved
    in one CG step, X_j = W / (sigma + tau_j)."""
    n, sigma = 50, 0.3
    G = pg.Graph(n, np.zeros((0, 2), dtype=np.int64), sigma)
    op = pg.Op(G)
    rng = np.random.default_rng(0)
    X = torch.from_numpy(rng.standard_normal((n, 6))).cuda()
    Y = op(X, 0.2)
    assert torch.equal(Y, (sigma + 0.2) * X)
    W = rng.standard_normal((n, 5))
    W[:, 2] = 0.0
    tau = np.array([0.0, 0.5, 1.0])
    act = np.arange(0, n, 3)
    Xs, h, it = pg.msolve(op, torch.from_numpy(W).cuda(), tau, act, 1e-12, track=True)
    assert it[2] == 0 and np.all(it[[0, 1, 3, 4]] == 1)
    want = W[act][None, :, :] / (sigma + tau)[:, None, None]
    assert np.max(np.abs(Xs.cpu().numpy() - want)) <= 1e-14 * np.max(np.abs(want))


def test_edgeless_graph_greedy_matches_oracle_and_closed_form():
    """All candidates join isolated vertices: the FP64 HIP greedy equals the oracle (sequence,
    counters, marginals), and while a pair touches no added edge its marginal is f(sigma) -
    f(sigma + 2) with f the pole rule."""
    n, sigma, s = 40, 0.3, 0.5
    e = np.zeros((0, 2), dtype=np.int64)
    cand = graphs.candidates(n, e, 30, 2)
    Go, G = O.Graph(n, e, sigma), pg.Graph(n, e, sigma)
    tau, om, _ = pg.fixed_rule(s, 8, sigma, G.Lam(5))
    sk = O.sketches(Go, cand, 8, 4, kind='rademacher')
    kw = dict(q=8, B=4, tol=1e-10, sketch=sk, poles=(tau, om))
    ro = O.polar(Go, cand, 5, s, arith='cr', **kw)
    rg = pg.polar(G, cand, 5, s, **kw)
    assert rg['sequence'] == ro['sequence'] and (rg['spmv'], rg['rhs']) == (ro['spmv'], ro['rhs'])
    assert np.max(np.abs(np.subtract(rg['marginals'], ro['marginals']))) <= 1e-13 * np.max(ro['marginals'])
    closed = float(np.sum(om / (sigma + tau)) - np.sum(om / (sigma + 2 + tau)))
    touched = set()
    for r, c in enumerate(rg['sequence']):
        u, v = (int(x) for x in cand[c])
        if u not in touched and v not in touched:
            assert abs(rg['marginals'][r] - closed) <= 1e-13 * closed
        touched |= {u, v}


def test_history_entry_points_check_arguments():
    """polar_history_rows / polar_history_cols: no POLAR_HISTORY flag, a workspace smaller than the
    history's, no iterations kept, more than 4 columns per call, no rows."""
    lib = L.require_gpu()
    op, s, dt, n = _small_op('fp32')
    R, p, nA, tcap = 4, 3, n, 32
    fl = L.POLAR_HISTORY | L.POLAR_BLOCK8
    need = lib.polar_msolve_history_ws_bytes(s, R, p, nA, fl, tcap)
    assert need > 0
    ws = torch.zeros(need, dtype=torch.uint8, device='cuda')
    rows = torch.zeros((2, R), dtype=torch.int32, device='cuda')
    E = torch.empty((p, 2, R), dtype=dt, device='cuda')
    cols = torch.zeros(5, dtype=torch.int32, device='cuda')
    act = torch.arange(n, dtype=torch.int32, device='cuda')
    X = torch.empty((5, p, nA), dtype=dt, device='cuda')
    sp = L.stream_ptr()
    rows_call = lambda f, nb, t, K: lib.polar_history_rows(s, R, p, nA, f, t, L.ptr(ws), nb, K, L.ptr(rows), L.ptr(E), sp)
    _expect_arg(rows_call(L.POLAR_BLOCK8, need, tcap, 2), 'without POLAR_HISTORY')
    _expect_arg(rows_call(fl, need - 256, tcap, 2), 'too small')
    _expect_arg(rows_call(fl, need, 0, 2), 'bad history arguments')
    _expect_arg(rows_call(fl, need, tcap, 0), 'bad history_rows arguments')
    _expect_arg(lib.polar_history_cols(s, R, p, nA, fl, tcap, L.ptr(ws), need, 5, L.ptr(cols), L.ptr(act), L.ptr(X), sp),
                'at most 4 columns')
    _expect_arg(lib.polar_history_cols(s, R, p, nA, fl, tcap, L.ptr(ws), need, 2, None, L.ptr(act), L.ptr(X), sp),
                'bad history_cols arguments')
