import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')
    config.addinivalue_line('markers', 'slow: longer CPU oracle runs')


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope='session')
def gpu():
    if not gpu_available():
        pytest.skip('no GPU visible')
    from sosfuse2_amd import hip
    hip.lib()  # fails loudly when the library is missing
    return True
