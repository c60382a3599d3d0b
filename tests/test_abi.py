"""CPU checks of the C ABI: the built library loads and exports every symbol
that include/jgcm.h declares, and the ctypes table covers them all.
No compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
           if f.endswith(".h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(jgcm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "jgcm_trans_grid_to_spherical" in names
    assert "jgcm_trans_spherical_to_grid" in names
    assert len(names) > 20


def test_library_exports_declared_symbols(jgcm):
    L = jgcm.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header(jgcm):
    from jgcm_amd import _lib
    missing = sorted(declared_symbols() - set(_lib.SIGNATURES))
    assert not missing, missing


def test_no_gpu_reports_error_not_crash(jgcm):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    n = ctypes.c_int(-1)
    rc = jgcm.lib().jgcm_device_count(ctypes.byref(n))
    assert rc != 0
    assert jgcm.lib().jgcm_last_error()
