import contextlib
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


@contextlib.contextmanager
def fork_all_rng():
    """Save and restore the python, numpy and torch RNG states."""
    py_state = random.getstate()
    np_state = np.random.get_state()
    try:
        with torch.random.fork_rng():
            yield
    finally:
        np.random.set_state(np_state)
        random.setstate(py_state)


@pytest.fixture
def restore_rng_state():
    with fork_all_rng():
        yield


@pytest.fixture(scope="session")
def gpu_device():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
