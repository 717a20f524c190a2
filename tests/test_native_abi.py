"""The C-ABI library loads on a machine without a GPU and exports every symbol the headers
under include/ declare (no compute calls here)."""
import glob
import os
import re

from snarkrs_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(g16(?:hip|zkey|wtns|synth)_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = native.load()
    for h in sorted(os.path.basename(p) for p in glob.glob(os.path.join(ROOT, "include", "*.h"))):
        missing = [s for s in _declared(h) if not hasattr(lib, s)]
        assert not missing, (h, missing)


def test_binding_declares_every_g16hip_symbol():
    assert sorted(native.EXPORTS) == _declared("g16hip.h")
