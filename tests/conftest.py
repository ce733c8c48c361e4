import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def line_model():
    from magritte_amd import models

    return models.model_line_1d()


@pytest.fixture(scope="session")
def small_wind():
    from magritte_amd import models

    return models.model_wind(n_edge=12, seed=models.seed_for(1, 7))


@pytest.fixture(scope="session")
def wind32():
    from magritte_amd import models

    return models.model_wind(n_edge=32)
