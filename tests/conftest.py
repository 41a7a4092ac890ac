"""pytest configuration: the `gpu` marker (tests that need an MI355X) and shared fixtures."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

DECKS = ROOT / "tests" / "golden" / "decks"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def lib():
    from openpeec_amd import load_library
    return load_library(build_if_missing=True)


def deck_text(name: str) -> str:
    return (DECKS / name).read_text()
