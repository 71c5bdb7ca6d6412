import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblpm_bve.so on cuda:0)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def omesh(oracle_mod):
    cache = {}

    def get(level):
        if level not in cache:
            cache[level] = oracle_mod.Mesh(level)
        return cache[level]
    return get


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import lpmkokkos_amd as L

    made = []

    def make():
        c = L.Context(0)
        made.append(c)
        return c
    yield make
    for c in made:
        c.close()
