import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden_cases():
    out = []
    for name in sorted(os.listdir(GOLDEN)):
        p = os.path.join(GOLDEN, name, "case.json")
        if os.path.exists(p):
            with open(p) as f:
                out.append((name, json.load(f)))
    return out


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.build()
    return O
