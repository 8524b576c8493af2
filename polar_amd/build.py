"""Builds the gfx950 HIP library ``polar_amd/libpolar_amd.so`` in-tree with hipcc.

    python -m polar_amd.build          # or: python polar_amd/build.py

The library is plain C-ABI (include/polar_amd.h); Python binds it with ctypes (polar_amd/_lib.py).
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = [HERE / 'csrc' / 'krylov.hip', HERE / 'csrc' / 'greedy.hip']
OUT = HERE / 'libpolar_amd.so'
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = ['-O3', '--offload-arch=gfx950', '-fPIC', '-shared', '-std=c++17', '-ffp-contract=off',
         # packed-FP32 ops stay off in every HIP build of this project
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', str(HERE.parent / 'include'), '-Wno-unused-command-line-argument']


def build(verbose=True):
    tmp = OUT.with_suffix('.so.tmp')
    cmd = [HIPCC, *FLAGS, *map(str, SRC), '-o', str(tmp)]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    build()
    sys.exit(0)
