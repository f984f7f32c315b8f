import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "voss-prime-gaps_v1_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: CPU test that runs for tens of seconds")


def _ensure_oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)
    return O


@pytest.fixture(scope="session")
def oracle():
    return _ensure_oracle()


@pytest.fixture(scope="session")
def engine():
    import vossgpu
    print(f"[gpu] loading {vossgpu.LIB_PATH}", flush=True)
    e = vossgpu.Engine(0)
    print(f"[gpu] device {e.name} cus={e.num_cus}", flush=True)
    yield e
    e.close()
