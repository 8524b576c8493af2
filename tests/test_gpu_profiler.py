"""The live event profiler (polar_prof_set / polar_prof_get) that bench.py reads: its modes count
the launches they time, class 7 counts every iteration, and profiling changes no result."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

from polar_amd import _lib as L  # noqa: E402
from polar_amd import graphs  # noqa: E402
from polar_amd import polar_gpu as pg  # noqa: E402


def _read(lib, cls):
    ms, la, by = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    L.check(lib.polar_prof_get(cls, ctypes.byref(ms), ctypes.byref(la), ctypes.byref(by)), 'prof_get')
    return ms.value, la.value, by.value


def _solve(G, W, tau, act):
    X, h, it = pg.msolve(pg.Op(G, precision='fp32'), W, tau, act, 1e-6, track=True)
    return X.cpu(), h.cpu(), it


def test_profiler_modes_count_and_change_nothing():
    lib = L.require_gpu()
    n = 3000
    G = pg.Graph(n, graphs.chung_lu(n, 2.3, 8.0, 4), 0.2)
    rng = np.random.default_rng(2)
    W = torch.from_numpy(rng.standard_normal((n, 8))).float().cuda()
    tau = np.array([0.0, 0.3, 2.0])
    act = np.arange(0, n, 5)
    ref = _solve(G, W, tau, act)
    iters = int(ref[2].max())
    assert iters > 16
    got = {}
    try:
        for mode in (1, 2, 3):
            L.check(lib.polar_prof_set(mode), 'prof_set')
            out = _solve(G, W, tau, act)
            assert np.array_equal(out[2], ref[2]) and torch.equal(out[0], ref[0]) and torch.equal(out[1], ref[1])
            got[mode] = [_read(lib, c) for c in range(8)]
    finally:
        lib.polar_prof_set(0)
    # every iteration is counted in class 7; modes 1 and 2 time all of them
    for mode in (1, 2, 3):
        assert got[mode][7][1] == iters
    assert got[1][0][1] == iters and got[2][0][1] == iters
    assert all(got[1][c][0] > 0 for c in range(5))
    assert got[2][1][0] == 0.0                       # mode 2 times the operator only
    # mode 3: one iteration per batch of 8, at position (batch % 8) within the batch
    sampled = sum(1 for b in range((iters + 7) // 8) if 8 * b + b % 8 < iters)
    assert got[3][0][1] == sampled and got[3][0][0] > 0
    per_launch = got[1][0][2] / got[1][0][1]
    assert got[3][0][2] / sampled == pytest.approx(per_launch, rel=0.5)
    assert lib.polar_prof_set(4) == L.POLAR_E_ARG
    assert lib.polar_prof_get(8, None, None, None) == L.POLAR_E_ARG
