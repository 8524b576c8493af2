"""ctypes binding of libpolar_amd.so (include/polar_amd.h).

The product path has no CPU fallback: importing a compute entry point without the built library,
or calling it without a ROCm GPU, raises.  Tensors are passed as raw device pointers.
"""
import ctypes
import os
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / 'libpolar_amd.so'

POLAR_F64, POLAR_F32 = 0, 1
POLAR_TRACK, POLAR_TRACE, POLAR_NOFREEZE, POLAR_GRAPH, POLAR_BLOCK8, POLAR_REPLACE = 1, 2, 4, 8, 16, 32
POLAR_HISTORY = 64
ABI_VERSION = 5
POLAR_E_ITERCAP, POLAR_E_ARG = -2, -3
CHUNK = 64  # POLAR_CHUNK: default long-row chunk length

EXPORTS = [
    'polar_abi_version', 'polar_last_error', 'polar_spmm_ws_bytes', 'polar_spmm',
    'polar_msolve_ws_bytes', 'polar_msolve', 'polar_pair_sqnorm', 'polar_proxy',
    'polar_select_ws_bytes', 'polar_select_topb', 'polar_pair_rhs', 'polar_rerank_pick',
    'polar_maint_prep', 'polar_maint_rank1', 'polar_maint_cache', 'polar_rows_axpy',
    'polar_rademacher', 'polar_edge_sketch', 'polar_prof_set', 'polar_prof_get', 'polar_pair_wsum',
    'polar_msolve_batched_ws_bytes', 'polar_msolve_batched', 'polar_csr_spmm', 'polar_release',
    'polar_spmm_static', 'polar_spmm_static_max_bytes', 'polar_msolve_history_ws_bytes', 'polar_history_rows',
    'polar_history_cols',
]

vp = ctypes.c_void_p
i64 = ctypes.c_int64
i32 = ctypes.c_int32
f64 = ctypes.c_double
u64 = ctypes.c_uint64


class PolarOp(ctypes.Structure):
    _fields_ = [
        ('n', i64), ('dtype', i32), ('weighted', i32),
        ('indptr', vp), ('indices', vp), ('weights', vp), ('diag', vp),
        ('n_long', i32), ('n_extra', i32), ('n_partial', i32), ('chunk', i32),
        ('long_rows', vp), ('long_base', vp), ('extra', vp),
        ('n_patch', i32), ('cpl_max', i32),
        ('patch_rows', vp), ('patch_ptr', vp), ('patch_cols', vp),
        ('grid_max', i32), ('run_rows', i32), ('patch_grp', vp),
    ]


class PolarSegs(ctypes.Structure):
    _fields_ = [('S', i32), ('G', i32), ('rows', vp), ('extra', vp), ('longs', vp), ('active', vp),
                ('col_group', vp)]


OPP = ctypes.POINTER(PolarOp)
SEGP = ctypes.POINTER(PolarSegs)

_SIGS = {
    'polar_abi_version': (i32, []),
    'polar_last_error': (ctypes.c_char_p, []),
    'polar_spmm_ws_bytes': (i64, [OPP, i64]),
    'polar_spmm': (i32, [OPP, f64, vp, i64, vp, i64, i64, vp, i64, vp]),
    'polar_csr_spmm': (i32, [i32, i64, i64, vp, vp, vp, vp, i64, vp, i64, vp]),
    'polar_release': (i32, []),
    'polar_spmm_static': (i32, [OPP, f64, vp, i64, vp, i64, i64, vp]),
    'polar_spmm_static_max_bytes': (i64, []),
    'polar_msolve_ws_bytes': (i64, [OPP, i64, i32, i64, i32]),
    'polar_msolve': (i32, [OPP, vp, i64, i64, vp, i32, vp, i64, f64, i64, i32, vp, i64, i64, vp, vp,
                           vp, vp, i64, vp]),
    'polar_pair_sqnorm': (i32, [i32, vp, i64, i64, i32, i64, vp, vp, i64, vp, vp]),
    'polar_proxy': (i32, [i32, vp, vp, vp, vp, i32, i64, f64, vp, vp, vp]),
    'polar_select_ws_bytes': (i64, [i64]),
    'polar_select_topb': (i32, [vp, i64, i32, vp, vp, vp, i64, vp]),
    'polar_pair_rhs': (i32, [i32, i64, i32, vp, vp, vp, vp]),
    'polar_rerank_pick': (i32, [i32, i32, vp, i64, i64, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp]),
    'polar_maint_prep': (i32, [i32, vp, vp, i64, i64, i32, i64, vp, i64, i64, i32, i32, vp, vp, vp]),
    'polar_maint_rank1': (i32, [i32, vp, vp, i64, i64, i32, i64, i64, vp, i64, i64, vp, vp, vp, vp, vp]),
    'polar_maint_cache': (i32, [i32, i64, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp]),
    'polar_rows_axpy': (i32, [i32, vp, i64, i64, i32, i32, vp, vp]),
    'polar_rademacher': (i32, [i32, u64, i32, vp, i64, i64, f64, vp, i64, vp]),
    'polar_edge_sketch': (i32, [OPP, u64, i64, f64, vp, i64, vp]),
    'polar_prof_set': (i32, [i32]),
    'polar_pair_wsum': (i32, [i32, vp, i64, i64, i32, i64, vp, vp, i64, vp, vp, i32, vp]),
    'polar_prof_get': (i32, [i32, vp, vp, vp]),
    'polar_msolve_batched_ws_bytes': (i64, [OPP, SEGP, i64, i32, i32]),
    'polar_msolve_batched': (i32, [OPP, SEGP, vp, i64, i64, vp, i32, vp, f64, i64, i32, vp, i64, i64, vp, vp,
                                   vp, vp, i64, vp]),
    'polar_msolve_history_ws_bytes': (i64, [OPP, i64, i32, i64, i32, i64]),
    'polar_history_rows': (i32, [OPP, i64, i32, i64, i32, i64, vp, i64, i32, vp, vp, vp]),
    'polar_history_cols': (i32, [OPP, i64, i32, i64, i32, i64, vp, i64, i32, vp, vp, vp, vp]),
}

_lib = None


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f'{LIB_PATH} is not built: run `python -m polar_amd.build` '
                           '(there is no CPU fallback)')
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.polar_abi_version() != ABI_VERSION:
        raise RuntimeError('libpolar_amd.so ABI mismatch')
    _lib = lib
    return lib


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError('polar_amd needs a ROCm GPU (torch.cuda.is_available() is False); '
                           'there is no CPU fallback')
    return load()


class PolarError(RuntimeError):
    pass


def check(rc, what=''):
    if rc == 0:
        return
    msg = _lib.polar_last_error().decode(errors='replace')
    if rc == POLAR_E_ITERCAP:
        raise RuntimeError('multi-shift CG iteration cap')
    raise PolarError(f'{what}: status {rc}: {msg}')


def ptr(t):
    """Raw device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def dtype_code(dtype):
    if dtype == torch.float64:
        return POLAR_F64
    if dtype == torch.float32:
        return POLAR_F32
    raise ValueError(f'unsupported dtype {dtype}')


def lib():
    return load()


os.environ.setdefault('HSA_ENABLE_IPC_MODE_LEGACY', '0')
