"""Shared fixtures.  CPU tests drive the oracle library (test infrastructure,
oracle/libcliquid_oracle.so); GPU tests (marked `gpu`) drive the HIP library
and compare its results with the oracle's on the same seeded inputs."""

from __future__ import annotations

import ctypes
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from continuum_liquid_amd import runtime  # noqa: E402

ORACLE_LIBRARY = ROOT / "oracle" / "libcliquid_oracle.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


def load_oracle() -> ctypes.CDLL:
    if not ORACLE_LIBRARY.exists():
        raise RuntimeError(f"{ORACLE_LIBRARY} is missing: run `make -C oracle` or __graft_entry__.build()")
    return runtime.bind_library(ctypes.CDLL(str(ORACLE_LIBRARY)))


@pytest.fixture(scope="session")
def oracle_lib():
    return load_oracle()


@pytest.fixture(scope="session")
def hip_lib():
    lib = runtime.load_library()  # raises when the HIP library is missing
    if runtime.device_count(lib) < 1:
        pytest.fail("no HIP device visible to a GPU test")
    return lib
