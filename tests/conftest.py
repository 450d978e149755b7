import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
for p in (ROOT, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    if not os.path.exists(os.path.join(ORACLE, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", ORACLE], check=True)


def gpu_case(msg):
    """Print and flush the name of a GPU case before it runs."""
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="session")
def ctx():
    from asend3d_amd import api
    c = api.Context(0)
    yield c
    c.close()
