"""Shared test setup: the `gpu` marker, import paths and helpers."""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libvoss.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def voss():
    so = ROOT / "voss_amd" / "libvoss.so"
    if not so.exists():
        pytest.fail("libvoss.so is not built: run `python voss_amd/build.py` first")
    import voss_amd
    return voss_amd
