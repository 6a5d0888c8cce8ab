import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "wukong-cube_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP path)")


def _ensure_built():
    import wukong_gpu
    import oracle
    if not os.path.exists(wukong_gpu.LIB_PATH):
        from wukong_gpu import _build
        _build.build()
    if not os.path.exists(oracle.LIB_PATH):
        from oracle import build
        build.build()


_ensure_built()


@pytest.fixture(autouse=True)
def _announce_gpu_test(request):
    """Name each GPU test on stdout before it touches the device."""
    if request.node.get_closest_marker("gpu"):
        print("\n[gpu-test] start %s" % request.node.nodeid, flush=True)
    yield


@pytest.fixture(scope="session")
def univ10():
    """Config-1 graph: university generator, seed 1, 10 universities."""
    import wukong_gpu as wg
    return wg.gen_university(1, 10)


@pytest.fixture(scope="session")
def univ10_oracle(univ10):
    import oracle
    return oracle.Store(univ10)


@pytest.fixture(scope="session")
def univ10_store(univ10):
    import wukong_gpu as wg
    return wg.Store(univ10, pinned=True)


@pytest.fixture(scope="session")
def univ10_dev(univ10_store):
    import wukong_gpu as wg
    dev = wg.Device(univ10_store, table_budget=1 << 30)
    dev.load_all()
    yield dev
    dev.close()
