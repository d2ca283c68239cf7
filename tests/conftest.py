import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "cfd-with-cuda_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE = "/root/reference"
INP_FIXTURE = os.path.join(REFERENCE, "inputFiles/lidDrivenCavity/lidDrivenCavity_NE27000.inp")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def oracle_lib():
    import fsx_oracle
    if not os.path.exists(fsx_oracle._LIB_PATH):
        fsx_oracle.build()
    return fsx_oracle


@pytest.fixture(scope="session")
def fsx_lib():
    import fsx
    fsx.lib()
    return fsx
