"""Zero-allocation SpMM (polar_spmm_static): the long-row chunk partials live in the library's static
device scratch instead of a caller workspace.  Called through the C-ABI, its output equals
polar_spmm's with a workspace (the same kernels) bit for bit and matches SciPy, on two calls in a
row -- the second with a smaller input (fewer rows, fewer long-row chunks, fewer columns), so the
scratch still holds the first call's partials when the second starts; an operator whose partials
exceed the static maximum is rejected on the host before any launch."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from polar_amd import _lib as L  # noqa: E402
from polar_amd import graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _hub_graph(n, seed):
    """Power-law graph plus two hubs longer than several 64-entry chunks (long-row partials)."""
    e = graphs.chung_lu(n, 2.3, 8.0, seed)
    hubs = np.array([[0, j] for j in range(1, min(n, 900))] + [[7, j] for j in range(300, min(n, 1000))])
    return graphs.simple_edges(np.concatenate([e, hubs]), n)


def _products(lib, G, pairs, R, precision, seed):
    op = pg.Op(G, precision=precision)
    for a, b in pairs:
        op.add(a, b)
    s = op.c_op()
    dt = op.dtype
    assert s.n_partial > 0
    X = torch.from_numpy(np.random.default_rng(seed).standard_normal((G.n, R))).to('cuda', dt)
    shift = 0.0375
    Ys = torch.full_like(X, float('nan'))
    L.check(lib.polar_spmm_static(s, shift, L.ptr(X), R, L.ptr(Ys), R, R, L.stream_ptr()), 'polar_spmm_static')
    ws = torch.empty(lib.polar_spmm_ws_bytes(s, R), dtype=torch.uint8, device='cuda')
    Yw = torch.full_like(X, float('nan'))
    L.check(lib.polar_spmm(s, shift, L.ptr(X), R, L.ptr(Yw), R, R, L.ptr(ws), ws.numel(), L.stream_ptr()),
            'polar_spmm')
    torch.cuda.synchronize()
    M = G.M0.get().tolil()
    for a, b in pairs:
        M[a, a] += 1.0
        M[b, b] += 1.0
        M[a, b] -= 1.0
        M[b, a] -= 1.0
    M = M.tocsr()
    Xh = X.double().cpu().numpy()
    ref = M @ Xh + shift * Xh
    scale = np.abs(M).sum(1).A * np.abs(Xh).max()
    return Ys, Yw, ref, scale


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_static_spmm_equals_workspace_spmm_twice(precision):
    lib = L.require_gpu()
    tol = 1e-12 if precision == 'fp64' else 2e-6
    calls = ((3000, 64, [(3, 2500), (0, 2999)], 1), (1100, 24, [(5, 9)], 2))   # the second is smaller
    for n, R, pairs, seed in calls:
        G = pg.Graph(n, _hub_graph(n, seed), 0.2)
        Ys, Yw, ref, scale = _products(lib, G, pairs, R, precision, seed)
        assert torch.equal(Ys, Yw), (n, R)
        assert np.max(np.abs(Ys.double().cpu().numpy() - ref) / scale) < tol, (n, R)


def test_static_spmm_rejects_partials_beyond_the_maximum():
    lib = L.load()
    mx = lib.polar_spmm_static_max_bytes()
    assert mx == 32 << 20
    ip = np.zeros(5, dtype=np.int32)
    s = L.PolarOp()
    s.n = 4
    s.dtype = L.POLAR_F64
    s.indptr = ip.ctypes.data_as(L.vp)
    s.diag = ip.ctypes.data_as(L.vp)
    s.chunk = L.CHUNK
    s.n_partial = mx // (8 * 64) + 1          # one chunk row more than fits at R=64
    assert lib.polar_spmm_ws_bytes(s, 64) > mx
    rc = lib.polar_spmm_static(s, 0.0, None, 64, None, 64, 64, None)
    assert rc == L.POLAR_E_ARG and b'static maximum' in lib.polar_last_error()
