"""Shared pytest setup: the `gpu` marker and loaders for the engine and the oracle."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_jsonl(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return [json.loads(line) for line in f if line.strip()]


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def engine():
    import perft_gpu_2026_amd as pg
    pg.lib()
    return pg


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(engine):
    """Initialised GPU engine (fails loudly if the HIP library cannot run)."""
    print("[gpu] pg_init", flush=True)
    engine.init(arena_bytes=32 << 30)
    yield engine
