"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import aperi_hip as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names.update(re.findall(r"^\s*(?:int|const char \*)\s*(aphip_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(A.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) <= set(A.EXPORTED)


def test_library_is_gfx950_code():
    data = open(A.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_error_path_without_context():
    L = A.load_library()
    assert L.aphip_set_mesh(None, 0, 0, None, 0, None) != 0
    assert b"NULL" in L.aphip_last_error() or len(L.aphip_last_error()) > 0
