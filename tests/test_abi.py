"""The C ABI library loads and exports every entry point include/ash3d_tr.h
declares (no device call is made, so this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ash3d_tr_\w+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"ash3d_tr_create", "ash3d_tr_step", "ash3d_tr_advect_x", "ash3d_tr_diffuse_dir"} <= names


def test_library_exports_every_declared_symbol(pkg):
    import importlib
    tr = importlib.import_module("volcano-ash3d_amd.transport")
    lib = ctypes.CDLL(tr.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes mirror covers the same set
    assert set(tr.EXPORTS) == declared_symbols()


def test_build_info_without_device(pkg):
    import importlib
    tr = importlib.import_module("volcano-ash3d_amd.transport")
    assert "gfx950" in tr.build_info()
