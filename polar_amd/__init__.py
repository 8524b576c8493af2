"""polar_amd: MI355X-native pole-shared Krylov core of POLAR (fractional spectral link selection).

Drop-in modules mirroring the reference's realworld/ layout:
  polar_amd.polar_gpu  -- Graph, Op, msolve, sketches, polar, true_marginals, ... (GPU)
  polar_amd.spmm       -- spmm(A, X)
  polar_amd.graphs     -- synthetic generators and candidate pools (host)
  polar_amd.run_real   -- `time` / `effect` CLI
Compute runs in libpolar_amd.so (hand-written gfx950 HIP kernels, C-ABI in include/polar_amd.h).
"""
__version__ = '0.1.0'
