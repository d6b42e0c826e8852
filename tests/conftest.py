import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def data_dir():
    return os.path.join(ROOT, "tests", "golden", "data")


@pytest.fixture(scope="session")
def cu01(data_dir):
    return os.path.join(data_dir, "Cu01.eam.alloy")


@pytest.fixture(scope="session")
def synth_setfl(tmp_path_factory):
    from fmd_amd import workloads
    p = str(tmp_path_factory.mktemp("eam") / "Cu_synth.eam.alloy")
    workloads.write_setfl(p, nr=5000, nrho=5000)
    return p


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
