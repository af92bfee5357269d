import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmgcd.so on the device)")


@pytest.fixture(scope="session")
def mgcd():
    import __graft_entry__ as ge
    return ge.package()


@pytest.fixture(scope="session")
def wl():
    import __graft_entry__ as ge
    return ge.workload()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib


_ctx_cache = {}


@pytest.fixture(scope="session")
def ctx_for(mgcd):
    """ctx_for(L) -> a Context with the L-bit table (one per width per session)."""
    def get(L):
        if L not in _ctx_cache:
            _ctx_cache[L] = mgcd.Context(L)
        return _ctx_cache[L]
    yield get
    for c in _ctx_cache.values():
        c.close()
    _ctx_cache.clear()
