import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)

GOLDEN = os.path.join(TESTS, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden_path(name):
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def twitter_large():
    with open(golden_path("twitter_sample_large_record.json"), "rb") as f:
        return f.read()


@pytest.fixture(scope="session")
def twitter_small():
    with open(golden_path("twitter_sample_small_records.json"), "rb") as f:
        return f.read()
