"""Shared pytest configuration: the `gpu` marker, the repo root on sys.path,
and a session-wide device context for GPU tests."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def ctx():
    from lucent_amd.runtime import Context
    c = Context(0)
    yield c
    c.close()
