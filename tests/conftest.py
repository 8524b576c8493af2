import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / 'tests')):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU and the built libpolar_amd.so')
    config.addinivalue_line('markers', 'slow: long CPU test')


def pytest_collection_modifyitems(config, items):
    gpu_ok = None
    for item in items:
        if 'gpu' in item.keywords:
            if gpu_ok is None:
                import torch
                gpu_ok = torch.cuda.is_available()
            if not gpu_ok:
                item.add_marker(pytest.mark.skip(reason='no ROCm GPU visible'))
