import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _print_case(item):
    # Name the case before any GPU work so a failing run says which one.
    if item.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu case] {item.nodeid}\n")
        sys.stdout.flush()


def pytest_runtest_setup(item):
    _print_case(item)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    return O
