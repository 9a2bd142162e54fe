import os
import sys

import pytest

try:  # one HIP runtime per process: torch binds it before the library loads
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "astro-accelerate_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def aa():
    import py_astro_accelerate as mod
    mod.load_library()
    return mod


@pytest.fixture(scope="session")
def oracle():
    import ddtr_oracle as mod
    mod.lib()
    return mod


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")
