import json
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")
CONFIGS = os.path.join(GOLDEN, "configs")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); run with -m gpu")


def config_path(name: str) -> str:
    return os.path.join(CONFIGS, name + ".json")


def golden(case: str) -> dict:
    with open(os.path.join(GOLDEN, f"ref_{case}.json")) as f:
        return json.load(f)


def golden_exists(case: str) -> bool:
    return os.path.exists(os.path.join(GOLDEN, f"ref_{case}.json"))


@pytest.fixture(scope="session")
def repo_root():
    return REPO
