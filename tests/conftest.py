"""Shared pytest setup: the `gpu` marker, import paths and size helpers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) and the built libmugpu.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    # the oracle's OpenMP loops on at most 16 host threads; only the summation
    # order of its dot products depends on the count
    O.set_threads(min(16, os.cpu_count() or 1))
    return O


@pytest.fixture(scope="session")
def capi():
    from muspectre_amd import _capi
    lib = _capi.load()
    _capi.check(lib.mugpu_init(0))
    return _capi
