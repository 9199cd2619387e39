import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def built():
    """The product library and the oracle, built in-tree."""
    from polymps_amd import build as b
    lib = b.LIB
    if not lib.exists():
        b.build()
    if not (ROOT / "oracle" / "liboracle_mps.so").exists():
        b.build_oracle()
    return lib
