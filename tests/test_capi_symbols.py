"""The C ABI library loads on the CPU and exports every symbol include/cdx.h declares."""
import ctypes
import os
import re

import pytest


def _declared(cdx):
    text = open(os.path.join(cdx.INCLUDE, "cdx.h")).read()
    return sorted(set(re.findall(r"\b(cdx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(cdx):
    assert os.path.exists(cdx.LIB_PATH), "libcdx.so not built"
    lib = ctypes.CDLL(cdx.LIB_PATH)
    declared = _declared(cdx)
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(cdx.EXPORTS) == declared


def test_no_device_is_reported_not_faked(cdx):
    """Without a GPU the library refuses to create a context (no CPU fallback)."""
    if cdx.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(cdx.CdxError):
        cdx.Context(0)


def test_status_strings(cdx):
    L = cdx.lib()
    assert L.cdx_version() >= 1
    for s in range(8):
        assert L.cdx_status_string(s)
