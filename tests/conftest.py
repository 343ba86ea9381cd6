import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gaseous-gusty-ball_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (run on the MI355X box)")
    config.addinivalue_line("markers", "slow: larger problem sizes")


@pytest.fixture(scope="session")
def ggb():
    import ggb as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def orc():
    import ggb_oracle as m
    return m


@pytest.fixture(scope="session")
def gpu(ggb):
    if ggb.device_init() != 0:
        pytest.fail("GPU test needs a HIP device")
    return True


@pytest.fixture(scope="session")
def ref(orc):
    """The compiled reference (oracle/_ref); one OpenMP thread so sums run in cell order."""
    if not orc.available("ref"):
        pytest.skip("reference build (oracle/_ref) not present")
    orc.set_omp_threads(1)
    yield orc
    orc.set_omp_threads(os.cpu_count() or 1)


def levels_for(ggb, orc, N):
    nl = orc.olib("f64").orc_mg_nlev(N)
    return [ggb.Grid(N >> l) for l in range(nl)]
