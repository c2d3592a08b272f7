import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def pytest_sessionstart(session):
    # the oracle is plain C; build it when the shared object is missing
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    # the HIP library before any test imports torch: the process then holds
    # one HIP runtime, the one the library is built against
    from tandemproject_amd import api
    try:
        api.load()
    except api.TPError:
        pass   # not built: the tests that need it say so


@pytest.fixture(autouse=True)
def _close_servers():
    """Close every request-loop server a test made, also when the test fails
    before its own close(): a server left open keeps its staging thread, its
    device memory and its page-locked result buffers for the rest of the
    session."""
    from tandemproject_amd import api
    made = []
    init = api.Server.__init__

    def tracking_init(self, *a, **k):
        made.append(self)
        init(self, *a, **k)

    api.Server.__init__ = tracking_init
    try:
        yield
    finally:
        api.Server.__init__ = init
        for s in made:
            s.close()
