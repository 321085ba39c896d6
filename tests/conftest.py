import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(TESTS, "golden")


@pytest.fixture(scope="session")
def expected_stdout(golden_dir):
    import json
    with open(os.path.join(golden_dir, "expected_stdout.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def example_puzzles(golden_dir):
    import oracle_bindings as o
    return {n: o.parse_puzzle(os.path.join(golden_dir, "puzzles", n + ".txt"))
            for n in ("3_3", "4_4", "5_5")}


@pytest.fixture(scope="session")
def gpu_context():
    import edge_matching_puzzle_amd as emp
    ctx = emp.Context(0)
    yield ctx
    ctx.close()
