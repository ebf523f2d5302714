import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libpq3d.so")


@pytest.fixture(scope="session")
def golden_td():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "td_golden.npz"), allow_pickle=False))


@pytest.fixture(scope="session", autouse=True)
def _private_tuning_cache(tmp_path_factory):
    """Simulations built by the tests tune their split-K widths into a cache of
    this session's own, not the user's."""
    if "PQ3D_TUNE_CACHE" not in os.environ:
        os.environ["PQ3D_TUNE_CACHE"] = str(tmp_path_factory.mktemp("tuning") / "tuning.json")
    yield
