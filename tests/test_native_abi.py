"""The C-ABI library loads on a machine without a GPU and exports every symbol
include/g16hip.h declares (no compute calls here)."""
import os
import re

from snarkrs_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "g16hip.h")).read()
    return sorted(set(re.findall(r"\b(g16hip_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = native.load()
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_declares_every_header_symbol():
    assert sorted(native.EXPORTS) == _declared()
