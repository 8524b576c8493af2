"""Batched form, host side: the union operator's segments and the batched workspace query."""
import ctypes

import numpy as np
import pytest

from polar_amd import _lib as L
from polar_amd import batched, graphs
from polar_amd import polar_gpu as pg


def _gs():
    return [pg.Graph(500, graphs.chung_lu(500, 2.3, 8.0, 1), 0.05),
            pg.Graph(3000, graphs.chung_lu(3000, 2.2, 10.0, 2), 0.1),
            pg.Graph(1200, graphs.barabasi_albert(1200, 4, 3), 0.2)]


def test_union_rows_equal_each_graph():
    Gs = _gs()
    U = batched.UnionGraph(Gs)
    for g, G in enumerate(Gs):
        r0, r1 = U.row0[g], U.row0[g + 1]
        ip = U._indptr[r0:r1 + 1] - U._indptr[r0]
        assert np.array_equal(ip, G._indptr)
        assert np.array_equal(U._indices[U._indptr[r0]:U._indptr[r1]] - r0, G._indices)
        assert np.array_equal(U.deg[r0:r1] + U.sigma[r0:r1], G.deg + G.sigma)   # same doubles


def test_union_segments_cover_each_schedule():
    Gs = _gs()
    U = batched.UnionGraph(Gs)
    rows, extra, longs, act = U.segments(np.array([0, 5, 9, 20]))
    sc = U.schedule()
    for g, G in enumerate(Gs):
        s = G.schedule()
        assert longs[g + 1] - longs[g] == len(s['long_rows'])
        assert extra[g + 1] - extra[g] == s['n_partial']
        assert np.array_equal(sc['long_rows'][longs[g]:longs[g + 1]] - rows[g], s['long_rows'])
    assert rows[-1] == U.n and extra[-1] == sc['n_partial'] and longs[-1] == len(sc['long_rows'])


def test_batched_workspace_query_through_ctypes():
    if not L.LIB_PATH.exists():
        pytest.skip('libpolar_amd.so not built')
    lib = L.load()
    Gs = _gs()
    U = batched.UnionGraph(Gs)
    sc = U.schedule()
    op = L.PolarOp()
    op.n, op.dtype, op.chunk = U.n, L.POLAR_F64, L.CHUNK
    op.indptr = op.diag = op.indices = ctypes.c_void_p(256)      # never dereferenced by the query
    op.n_long, op.n_extra, op.n_partial = len(sc['long_rows']), len(sc['extra']), sc['n_partial']
    op.long_rows = op.long_base = op.extra = ctypes.c_void_p(256)
    segs = batched.Segments(*U.segments(np.array([0, 5, 9, 20])))
    one = lib.polar_msolve_ws_bytes(op, 8, 6, 20, 0)
    many = lib.polar_msolve_batched_ws_bytes(op, segs.c, 8, 6, 0)
    assert one > 0 and many > one
    bad = batched.Segments(np.array([0, 5, 9, U.n - 1]), *U.segments(np.array([0, 5, 9, 20]))[1:])
    assert lib.polar_msolve_batched_ws_bytes(op, bad.c, 8, 6, 0) == -1
    op.cpl_max = 3
    assert lib.polar_msolve_ws_bytes(op, 8, 6, 20, 0) == -1
