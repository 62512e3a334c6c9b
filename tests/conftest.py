import random
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the C ABI)")


class RngSnapshot:
    """The state of every random generator a test may draw from: Python's
    ``random``, numpy's global generator, torch's CPU generator and, once
    CUDA is initialised, each device generator.  ``restore()`` puts them
    all back, so seeded tests neither see nor leave draws of other tests."""

    def __init__(self):
        self.python = random.getstate()
        self.numpy = np.random.get_state()
        self.torch_cpu = torch.get_rng_state()
        self.torch_cuda = (torch.cuda.get_rng_state_all()
                           if torch.cuda.is_available() and torch.cuda.is_initialized() else None)

    def restore(self) -> None:
        random.setstate(self.python)
        np.random.set_state(self.numpy)
        torch.set_rng_state(self.torch_cpu)
        if self.torch_cuda is not None:
            torch.cuda.set_rng_state_all(self.torch_cuda)


@pytest.fixture
def restore_rng_state(request):
    """Run the test on its own copy of the random streams."""
    snap = RngSnapshot()
    request.addfinalizer(snap.restore)
    return snap


@pytest.fixture(scope="session")
def gpu_device():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
