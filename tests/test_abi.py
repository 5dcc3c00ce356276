"""The C-ABI library loads and exports every symbol include/gnx.h declares.
No GPU calls are made here."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "3d-gross-neveu-model_amd", "lib", "libgnx.so")
HDR = os.path.join(ROOT, "include", "gnx.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gnx_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_and_exports():
    assert os.path.exists(LIB), "libgnx.so not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} missing from libgnx.so"
    assert lib.gnx_abi_version() == 3


def test_library_does_not_link_oracle():
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", LIB], capture_output=True, text=True).stdout
    assert "orc_" not in syms
