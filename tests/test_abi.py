"""The product library loads and exports every entry point include/mega.h
declares (no device call is made, so this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mega.h")
LIB = os.path.join(ROOT, "taylor_amd", "libmega_hip.so")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mega_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_step_interface():
    names = declared()
    for required in ("mega_create", "mega_search", "mega_shapes", "mega_interp", "mega_fint", "mega_constitute",
                     "mega_step", "mega_download", "mega_destroy"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("taylor_amd/libmega_hip.so is not built (python taylor_amd/build.py)")
    lib = C.CDLL(LIB)
    for name in declared():
        assert hasattr(lib, name), name


def test_status_strings_without_device():
    lib = C.CDLL(LIB)
    lib.mega_status_string.restype = C.c_char_p
    assert lib.mega_status_string(0) == b"ok"
    assert b"converge" in lib.mega_status_string(1)


def test_python_host_refuses_to_run_without_library(monkeypatch):
    from taylor_amd import mega
    monkeypatch.setattr(mega, "LIB_PATH", os.path.join(ROOT, "does_not_exist.so"))
    monkeypatch.setattr(mega, "_lib", None)
    with pytest.raises(RuntimeError):
        mega.load_library()
