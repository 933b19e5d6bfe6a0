"""Shared pytest configuration.

GPU tests carry @pytest.mark.gpu and run on a real MI355X; everything else is
CPU-only (the oracle, the host application layer, the C ABI's argument
checks and its exported symbols).
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding as ob

    ob.lib()
    return ob
