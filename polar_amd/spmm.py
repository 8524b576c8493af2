"""Drop-in for realworld/spmm.py: spmm(A, X) = M0 @ X on the GPU (HIP kernel polar_spmm)."""
from .polar_gpu import spmm  # noqa: F401
