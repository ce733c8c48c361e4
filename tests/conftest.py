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


@pytest.fixture(scope="session")
def vgrad_model():
    """1D spherical velocity-gradient benchmark model; max_width_fraction 0.10
    sends every segment through the erf (line-integrated) branch."""
    from magritte_amd import models

    ld = models.read_lamda(os.path.join(ROOT, "tests", "golden", "lamda_test.txt"), num=0)
    return models.model_velocity_gradient_1d(ld, max_width_fraction=0.10)


@pytest.fixture(scope="session")
def slab_model():
    from magritte_amd import models

    ld = models.read_lamda(os.path.join(ROOT, "tests", "golden", "lamda_test.txt"), num=0)
    return models.model_constant_slab(linedata=ld)


@pytest.fixture(scope="session")
def imager_model():
    """The reference's new-imager benchmark model (50 points, 4 rays, 10 quadrature bins)."""
    from magritte_amd import models

    ld = models.read_lamda(os.path.join(ROOT, "tests", "golden", "lamda_test.txt"), num=0)
    return models.model_velocity_gradient_1d(ld, npoints=50, nrays=4, nquads=10,
                                             max_width_fraction=0.5)
