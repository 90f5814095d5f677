"""Test configuration: the `gpu` marker and shared fixtures.

GPU tests (marked ``gpu``) run on the MI355X box and call the engine through the C ABI of
scout_amd/libscout_hip.so; every other test runs on the CPU.
"""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

GEN = os.path.join(REPO, "scout_amd", "bin", "scout_gen")
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libscout_hip.so)")


def run_gen(*args):
    """Runs scout_gen (the host generator) and raises on failure."""
    r = subprocess.run([GEN, *map(str, args)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"scout_gen {' '.join(map(str, args))} failed: {r.stderr}")
    return r.stdout


@pytest.fixture(scope="session")
def sample_dir():
    return os.path.join(GOLDEN, "sample")
