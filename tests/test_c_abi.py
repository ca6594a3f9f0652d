"""CPU checks of the C ABI: the library loads and exports every entry point include/afq.h
declares, and the Python mirror binds exactly those (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afq.h")
LIB = os.path.join(ROOT, "safire_amd", "libafq_hip.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(afq_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "afq_vbias" in names and "afq_prop_step" in names and len(names) >= 25


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_the_header():
    from safire_amd import afq
    assert sorted(afq.exported_symbols()) == declared()
