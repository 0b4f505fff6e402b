"""Shared pytest setup: the `gpu` marker and import paths for the package and the oracle."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "aeros-engine_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU reference runs (still part of the default suite)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    O.threads(min(8, __import__("os").cpu_count() or 1))
    return O
