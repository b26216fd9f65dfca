"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import Reference
    if not Reference.available():
        pytest.skip("compiled reference (oracle/_ref) not built here")
    return Reference()


@pytest.fixture(scope="session")
def ctx():
    from faspxx_amd import fx
    c = fx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def problems():
    """Parity-sized instances of the five configurations (cached per session)."""
    from faspxx_amd import problems as P
    cache = {}

    def get(cid, scale=None):
        key = (cid, scale)
        if key not in cache:
            cache[key] = P.config(cid, scale)
        return cache[key]
    return get


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
