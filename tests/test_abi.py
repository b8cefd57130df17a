"""The C-ABI library loads and exports every entry point its headers declare (CPU, no GPU calls)."""
import ctypes
import glob
import os
import re

import vecgeom_amd as vg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "vgm", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(vgm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    assert len(declared_symbols()) >= 15


def test_library_exports_all_declared_symbols():
    L = vg.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_last_error_is_callable_without_gpu():
    L = vg.lib()
    assert isinstance(L.vgm_last_error(), bytes)


def test_null_context_is_rejected_without_gpu():
    L = vg.lib()
    out = ctypes.c_int()
    assert L.vgm_box_create(None, 1.0, 1.0, 1.0, ctypes.byref(out)) != 0
