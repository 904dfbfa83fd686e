"""The C ABI library loads and exports every entry point include/casm_mi355x/mcgpu.h
declares (CPU only: no compute call is made)."""
import ctypes
import os
import re

import casmcode_monte_amd as cm

HEADER = os.path.join(cm.INCLUDE_DIR, "casm_mi355x", "mcgpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**(mcg_\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_declarations_match_binding_list():
    assert declared() == sorted(cm.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(cm.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name


def test_create_without_gpu_fails_loudly():
    """No CPU fallback: on a machine without a gfx950 device mcg_create errors."""
    import pytest
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(cm.McgError):
        cm.Run([dict(shape=(4, 4), temperature=1000.0, seed=1)])
