import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import lmref

    lmref.load()
    return lmref


@pytest.fixture(scope="session")
def gpu():
    """One lmgpu context for the GPU tests (fails loudly if the library or GPU is missing)."""
    from lightmetrica_amd.gpu import Device

    dev = Device(0)
    yield dev
    dev.close()


def flush_print(*a):
    print(*a, flush=True)
