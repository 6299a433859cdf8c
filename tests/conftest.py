import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# GPU files run in this order: the oracle-parity suites of the hot path come
# first, so that a failure in an addition (batches, serving, files) can never
# keep them from running under `pytest -x`.  Files not listed keep their
# place after the listed ones; the allocation-free entry point runs last.
_FIRST = [
    "test_gpu_rflu_parity.py",
    "test_gpu_picl_parity.py",
    "test_gpu_configs_parity.py",
    "test_gpu_ppiclf_f77.py",
    "test_gpu_fp32_mode.py",
    "test_gpu_flowio.py",
    "test_gpu_batch_serve.py",
]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def pytest_collection_modifyitems(session, config, items):
    def rank(item):
        name = os.path.basename(str(item.fspath))
        if name.startswith("test_zz_"):
            return len(_FIRST) + 1
        return _FIRST.index(name) if name in _FIRST else len(_FIRST)
    items.sort(key=rank)  # stable: the order inside a file is kept


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    # one process: the oracle's parallel loops may use the host's cores
    # (a pytest-xdist worker keeps one thread, beside the other workers)
    if "PYTEST_XDIST_WORKER" not in os.environ:
        oracle.set_threads(min(16, os.cpu_count() or 1))
    return oracle
