import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

# under pytest-xdist, share the cores between the workers' OpenMP pools (the
# oracle library reads this when it is first loaded, after this point)
if "PYTEST_XDIST_WORKER_COUNT" in os.environ:
    _n = max(1, (os.cpu_count() or 1) // int(os.environ["PYTEST_XDIST_WORKER_COUNT"]))
    os.environ["OMP_NUM_THREADS"] = str(_n)
for p in (ROOT / "oracle", ROOT / "qbox-public_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    try:
        oracle.lib()
    except RuntimeError:
        pytest.skip("oracle not built (run __graft_entry__.build())")
    return oracle
