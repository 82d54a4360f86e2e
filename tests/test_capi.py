"""Host-only checks of the C-ABI library: it loads, and exports every entry point that
include/stwo_hip.h declares (no compute call, no GPU needed)."""
import ctypes
import os
import re

import pytest

from hip_lib import SH, ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "stwo_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(stwo_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "stwo_hip_prove_poseidon" in names and "stwo_hip_commit_on_layer" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(SH.LIB_PATH):
        pytest.skip("libstwo_hip.so not built")
    lib = ctypes.CDLL(SH.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(SH.EXPORTS) == set(declared_functions())


def test_fnv1a_matches_reference_ids():
    # FNV-1a 32-bit (crates/constraint-framework/src/lib.rs fnv1a_eval_id_gen); standard test vector
    assert SH.fnv1a("") == 0x811C9DC5
    assert SH.fnv1a("a") == 0xE40C292C
