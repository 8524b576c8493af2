"""Drop-in for realworld/spmm.py: the SpMM Y = A X on the GPU.

* spmm(A, X) -- realworld/spmm.py:18-25.  A Graph, its M0 or an Op goes through the Laplacian
  operator kernel polar_spmm (value-free adjacency + diagonal + added-edge patch, the form every
  solve uses); any other CSR matrix through csr_spmm.
* csr_spmm(A, X) -- the general CSR kernel polar_csr_spmm (any values, diagonal stored), which
  replaces the reference's RawKernel csr_spmm (realworld/spmm.py:3-14).
"""
import scipy.sparse as sp
import torch

from . import _lib as L


def _csr_parts(A):
    """(rows, indptr, indices, data) of a CSR matrix: a torch sparse CSR tensor, a SciPy sparse
    matrix (converted to CSR on the host), or any object with indptr / indices / data / shape."""
    if isinstance(A, torch.Tensor):
        if A.layout != torch.sparse_csr:
            raise TypeError('spmm: a torch tensor A must have the sparse CSR layout')
        return A.shape[0], A.crow_indices(), A.col_indices(), A.values()
    if sp.issparse(A):
        A = A.tocsr()
    for att in ('indptr', 'indices', 'data', 'shape'):
        if not hasattr(A, att):
            raise TypeError(f'spmm: A has no {att!r}; pass a Graph, its M0, an Op or a CSR matrix')
    return A.shape[0], A.indptr, A.indices, A.data


def csr_spmm(A, X):
    """Y = A X for a general CSR matrix A (any values, diagonal stored) on the GPU: the kernel
    polar_csr_spmm, which replaces the reference's csr_spmm (realworld/spmm.py:3-14).  X: (m, R) or
    (m,), float64 or float32 (A's values are used in X's dtype); Y is (rows, R) (or 1-D)."""
    lib = L.require_gpu()
    dev = torch.device('cuda', torch.cuda.current_device())
    X = torch.as_tensor(X, device=dev)
    if X.dtype not in (torch.float64, torch.float32):
        X = X.to(torch.float64)
    one = X.ndim == 1
    X2 = (X[:, None] if one else X).contiguous()
    rows, ip, ix, va = _csr_parts(A)
    if hasattr(A, 'shape') and A.shape[1] != X2.shape[0]:
        raise ValueError(f'spmm: A is {tuple(A.shape)} but X has {X2.shape[0]} rows')
    ip = torch.as_tensor(ip, device=dev).to(torch.int32).contiguous()
    ix = torch.as_tensor(ix, device=dev).to(torch.int32).contiguous()
    va = torch.as_tensor(va, device=dev).to(X2.dtype).contiguous()
    R = X2.shape[1]
    Y = torch.empty((rows, R), dtype=X2.dtype, device=dev)
    if rows and R:
        L.check(lib.polar_csr_spmm(L.dtype_code(X2.dtype), rows, R, L.ptr(ip), L.ptr(ix), L.ptr(va), L.ptr(X2),
                                   X2.stride(0), L.ptr(Y), Y.stride(0), L.stream_ptr()), 'polar_csr_spmm')
    return Y[:, 0] if one else Y


def spmm(A, X):
    """Y = A X (realworld/spmm.py:18-25).  A Graph, its M0 (LaplacianCSR) or an Op runs through the
    Laplacian operator kernel (value-free adjacency + diagonal + added edges); any other CSR matrix
    (SciPy sparse, torch sparse CSR, or an object with indptr / indices / data / shape) through the
    general CSR kernel.  The reference sends 1-D X and R < 8 to cuSPARSE (`A @ X`); here every
    width runs the same kernels, which compute the same product."""
    from . import polar_gpu as pg
    if isinstance(A, pg.LaplacianCSR):
        A = A._G
    if isinstance(A, (pg.Graph, pg.Op)):
        op = A if isinstance(A, pg.Op) else pg.Op(A, precision=torch.as_tensor(X).dtype)
        return pg._apply(op, X, 0.0)
    return csr_spmm(A, X)
