"""Shared test configuration.

Markers: `gpu` tests need an MI355X and run the HIP path through the C ABI;
everything else runs on the CPU (oracle pinning, host logic, the library's
exported symbols).
"""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, os.path.join(REPO, "plp-gpu_amd"))
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def golden_path(*parts):
    return os.path.join(GOLDEN, *parts)


@pytest.fixture(scope="session")
def ref_runs():
    with open(golden_path("ref_runs.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def rng_vectors():
    with open(golden_path("rng_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def repr_vectors():
    with open(golden_path("repr_vectors.json")) as f:
        return json.load(f)
