import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "zinc-plus_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.load()
    o.set_threads(min(8, os.cpu_count() or 1))
    return o


@pytest.fixture(scope="session")
def zg():
    """The zipgpu package on cuda:0 (GPU tests only)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import zipgpu
    zipgpu.lib()
    assert zipgpu.lib().zp_init(0) == 0
    return zipgpu


@pytest.fixture(scope="session")
def dev():
    import torch
    return torch.device("cuda:0")
