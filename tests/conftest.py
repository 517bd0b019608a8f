"""Shared pytest configuration.

``gpu``-marked tests need an MI355X (HIP device); everything else runs on the
CPU.  The golden fixtures live in tests/golden (data copied from the
reference's peripy/test/data by tests/golden/collect_reference_fixtures.py).
"""
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = pathlib.Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs a HIP GPU (MI355X, gfx950)")


@pytest.fixture(scope="session")
def golden():
    return TESTS / "golden"
