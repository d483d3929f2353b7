"""CPU checks of the C ABI: the HIP library loads and exports exactly what
include/*.h declares.  No device calls."""
from __future__ import annotations

import glob
import os
import re

import pytest

from itk_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(mmi_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "mmi_evaluate" in names and "mmi_context_create" in names
    assert names == set(_native.SIGNATURES), names ^ set(_native.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("HIP library not built: run python -m itk_amd.build")
    L = _native.load()
    for name in _declared():
        assert hasattr(L, name), name


def test_last_error_is_a_string_before_any_call():
    L = _native.load()
    assert isinstance(L.mmi_last_error(), bytes)
