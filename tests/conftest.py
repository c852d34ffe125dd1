import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

# host preparations (grid geometry, sun table, ground pools, cell packs) of the configurations' full grids, built
# before any test starts the GPU runtime: their process pools are new interpreters, which a process holding the GPU
# runtime should not start, and the thread-pool fallback takes minutes at config1's size
PREPARED = {}


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')
    config.addinivalue_line('markers', 'slow: longer CPU oracle runs')


def _device_count():
    try:
        import torch

        return torch.cuda.device_count()  # counts devices without starting the runtime
    except Exception:
        return 0


@pytest.hookimpl(trylast=True)
def pytest_collection_modifyitems(session, config, items):
    needed = {name for it in items for name in ('config1_host', 'config0_host') if name in it.fixturenames}
    if not needed or _device_count() < 1:
        return
    from types import SimpleNamespace

    from sosfuse2_amd import api, atmosphere

    if 'config1_host' in needed:
        cfg = dict(api.CONFIG1, sun_scale=1)
        PREPARED['config1'] = api.Scope(cfg, atmosphere.config1_state(), SimpleNamespace(L=int(cfg['L'])))
    if 'config0_host' in needed:
        cfg = dict(api.CONFIG0, sun_scale=1)
        PREPARED['config0'] = api.Scope(cfg, atmosphere.config0_state(), SimpleNamespace(L=int(cfg['L'])))


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


def _scope(name, gpu):
    from sosfuse2_amd.kernel import Kernel

    scope = PREPARED.get(name)
    if scope is None:
        pytest.fail(f'the {name} host preparation was not built at collection')
    if not isinstance(scope.kernel, Kernel):
        scope.kernel = Kernel(backend='hip', L=scope.kernel.L)
    return scope


@pytest.fixture(scope='session')
def config1_host(gpu):
    """The config1 grid's scope (150 levels x 75 chi, L = 32, nmu 4, nphi 8, path_n 6) with a HIP kernel."""
    return _scope('config1', gpu)


@pytest.fixture(scope='session')
def config0_host(gpu):
    """The config0 grid's scope (60 levels x 10 chi, L = 16) with a HIP kernel."""
    return _scope('config0', gpu)
