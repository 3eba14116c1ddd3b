"""The C-ABI libraries load and export exactly what include/cliquid.h declares
(no compute calls: these run without a GPU)."""

from __future__ import annotations

import ctypes
import re

import pytest

from continuum_liquid_amd import runtime

from .conftest import ORACLE_LIBRARY, ROOT

HEADER = ROOT / "include" / "cliquid.h"


def declared_symbols() -> set[str]:
    text = HEADER.read_text()
    return set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(clq_\w+)\s*\(", text, re.M))


def test_header_and_binding_agree():
    assert declared_symbols() == set(runtime.EXPORTED_SYMBOLS)


@pytest.mark.parametrize("path", [runtime.HIP_LIBRARY, ORACLE_LIBRARY], ids=["hip", "oracle"])
def test_library_exports_every_symbol(path):
    assert path.exists(), f"{path} not built"
    lib = ctypes.CDLL(str(path))
    expected = declared_symbols()
    if path == ORACLE_LIBRARY:  # the oracle restates the step on host arrays; the network's CPU check is a
        expected -= set(runtime.GPU_ONLY_SYMBOLS)  # torch model
    missing = [name for name in expected if not hasattr(lib, name)]
    assert not missing, missing
    bound = runtime.bind_library(lib)
    assert bound.clq_abi_version() == runtime.ABI_VERSION


def test_hip_library_is_gfx950_only():
    data = runtime.HIP_LIBRARY.read_bytes()
    assert b"gfx950" in data
    for other in (b"gfx90a", b"gfx942", b"gfx1100"):
        assert other not in data


def test_load_library_fails_loudly_when_missing(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        runtime.load_library(tmp_path / "libcliquid_hip.so")
