"""Shared test setup: the `gpu` marker and the native libraries.

Tests load what is built; a library that is missing entirely is built once
here (the product with shocovox_amd.build, the checker with make)."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
TESTS = Path(__file__).resolve().parent
for p in (str(ROOT), str(TESTS)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    from shocovox_amd import _native
    if not _native.LIB_PATH.exists():
        from shocovox_amd import build
        build.build()
    import oracle
    if not oracle.LIB_PATH.exists():
        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")
