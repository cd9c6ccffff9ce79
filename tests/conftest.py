import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def ctx(pkg):
    return pkg.default_context()


@pytest.fixture(autouse=True)
def _svd_backend_override(request):
    """DMRJ_TEST_SVD_BACKEND=k runs every GPU test with SVD backend k on the
    default context (tests that choose a backend themselves still do)."""
    mode = os.environ.get("DMRJ_TEST_SVD_BACKEND")
    if mode is None or request.node.get_closest_marker("gpu") is None:
        yield
        return
    ctx = request.getfixturevalue("ctx")
    ctx.set_svd_backend(int(mode))
    yield
    ctx.set_svd_backend(0)
