"""The C-ABI library loads and exports every symbol that include/polar_amd.h declares (CPU: no
compute calls), and the ctypes struct matches the header's layout."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / 'include' / 'polar_amd.h').read_text()
    return sorted(set(re.findall(r'^\s*(?:int|int64_t|const char\*)\s+(polar_\w+)\s*\(', text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert 'polar_msolve' in syms and 'polar_spmm' in syms and len(syms) >= 18


def test_library_exports_every_declared_symbol():
    from polar_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip('libpolar_amd.so not built')
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(declared_symbols()) == set(_lib.EXPORTS)
    loaded = _lib.load()
    assert loaded.polar_abi_version() == _lib.ABI_VERSION == 5


def test_polar_op_struct_layout():
    from polar_amd._lib import PolarOp
    off = {name: getattr(PolarOp, name).offset for name, _ in PolarOp._fields_}
    assert off['indptr'] == 16 and off['diag'] == 40 and off['n_long'] == 48 and off['chunk'] == 60
    assert off['long_rows'] == 64 and off['n_patch'] == 88 and off['cpl_max'] == 92
    assert off['patch_cols'] == 112 and off['grid_max'] == 120 and off['patch_grp'] == 128
    assert ctypes.sizeof(PolarOp) == 136
    from polar_amd._lib import PolarSegs
    assert PolarSegs.rows.offset == 8 and PolarSegs.active.offset == 32 and PolarSegs.col_group.offset == 40
    assert ctypes.sizeof(PolarSegs) == 48


def test_product_path_has_no_cpu_fallback(monkeypatch):
    import torch
    from polar_amd import polar_gpu as pg
    if torch.cuda.is_available():
        pytest.skip('GPU present')
    with pytest.raises(RuntimeError):
        pg.msolve(None, [[1.0]], [1.0], [0], 1e-7)


def test_workspace_queries_through_ctypes():
    """Host-only entry points called with a real polar_op built on the host (no GPU needed)."""
    import numpy as np
    from polar_amd import _lib as L
    from polar_amd import graphs
    if not L.LIB_PATH.exists():
        pytest.skip('libpolar_amd.so not built')
    lib = L.load()
    n = 2000
    e = graphs.chung_lu(n, 2.3, 8.0, 1)
    hub = np.array([[0, j] for j in range(1, 700)])
    e = graphs.simple_edges(np.concatenate([e, hub]), n)
    ip = np.zeros(n + 1, dtype=np.int32)
    s = L.PolarOp()
    s.n = n
    s.dtype = L.POLAR_F64
    s.indptr = ip.ctypes.data_as(L.vp)
    s.diag = ip.ctypes.data_as(L.vp)
    s.n_partial = 3
    s.chunk = L.CHUNK
    s.n_long = 0
    assert lib.polar_spmm_ws_bytes(s, 16) >= 3 * 16 * 8
    ws = lib.polar_msolve_ws_bytes(s, 32, 24, 500, 0)
    assert ws >= 8 * (3 * n * 32 + 24 * 500 * 32)
    s.dtype = 7
    assert lib.polar_spmm_ws_bytes(s, 16) == -1
    assert lib.polar_select_ws_bytes(10 ** 6) > 0


def test_csr_spmm_and_release_host_checks():
    """polar_csr_spmm checks its arguments on the host before any launch (no GPU needed), and
    polar_release with nothing allocated is a no-op."""
    from polar_amd import _lib as L
    if not L.LIB_PATH.exists():
        pytest.skip('libpolar_amd.so not built')
    lib = L.load()
    dummy = ctypes.c_void_p(256)
    assert lib.polar_csr_spmm(7, 10, 4, dummy, dummy, dummy, dummy, 4, dummy, 4, None) == L.POLAR_E_ARG
    assert lib.polar_csr_spmm(L.POLAR_F64, 10, 4, dummy, dummy, dummy, dummy, 3, dummy, 4, None) == L.POLAR_E_ARG
    assert lib.polar_csr_spmm(L.POLAR_F32, 10, 0, dummy, dummy, dummy, dummy, 4, dummy, 4, None) == L.POLAR_E_ARG
    assert lib.polar_csr_spmm(L.POLAR_F64, 10, 4, None, dummy, dummy, dummy, 4, dummy, 4, None) == L.POLAR_E_ARG
    assert b'csr_spmm' in lib.polar_last_error()
    assert lib.polar_release() == 0
