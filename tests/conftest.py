import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libfvc.so")


def _gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def small_wl(oracle_mod):
    from frustumvolumecaching_amd import synthetic as S
    return S.small_plain(stepsize_setup=oracle_mod.stepsize_setup)


@pytest.fixture(scope="session")
def small_model(oracle_mod, small_wl):
    return oracle_mod.Model.from_workload(small_wl)


@pytest.fixture(scope="session")
def unbounded_wl(oracle_mod):
    from frustumvolumecaching_amd import synthetic as S
    return S.unbounded_1080p(width=240, height=136, stepsize_setup=oracle_mod.stepsize_setup)


@pytest.fixture(scope="session")
def unbounded_model(oracle_mod, unbounded_wl):
    return oracle_mod.Model.from_workload(unbounded_wl)
