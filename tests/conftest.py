import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")
REGISTRY = os.path.join(GOLDEN, "registry.bin")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    import cpu_oracle

    return cpu_oracle.Oracle(REGISTRY)


@pytest.fixture(scope="session")
def reference():
    import cpu_oracle

    if not cpu_oracle.Reference.available():
        pytest.skip("the reference mesher was not built here (oracle/_ref needs /root/reference)")
    return cpu_oracle.Reference()


@pytest.fixture(scope="session")
def ctx():
    from hypercraft_amd import hcm

    c = hcm.Context(pool_quads=1 << 24)
    yield c
    c.close()
