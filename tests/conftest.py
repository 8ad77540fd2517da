import os
import sys

import pytest

# torch first: its HIP runtime is the one every library of the package must bind to
import torch  # noqa: F401,E402

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libnsw6.so")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def grid():
    from nicam_kokkos_microphysics_amd.fields import synth_grid
    return synth_grid()
