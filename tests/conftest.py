import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuda-poker-solver_amd")
for p in (os.path.join(ROOT, "oracle"), PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU-only check")
    # The checker library is plain g++ output; build it when a fresh checkout
    # runs the CPU suite before anything else built it.
    lib = os.path.join(ROOT, "oracle", "bin", "liborc.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True,
                       stdout=subprocess.DEVNULL)


def card(text):
    return "cdhs".index(text[1].lower()) * 13 + "23456789TJQKA".index(text[0].upper())


def board_of(text):
    return [card(text[i:i + 2]) for i in range(0, len(text), 2)]


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")
