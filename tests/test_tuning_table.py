"""The launch-configuration table's reader (mmi_load_tuning_table): the
shipped table parses, malformed rows are refused with their line, and a
missing file is an error (no GPU call)."""
from __future__ import annotations

import ctypes
import os

import pytest

from itk_amd import _native as N

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "itk_amd", "tuning", "mi355x.tsv")


def _load(path):
    n = ctypes.c_int32(-1)
    rc = N.load().mmi_load_tuning_table(path.encode() if path else None, ctypes.byref(n))
    return rc, n.value, N.load().mmi_last_error().decode()


def test_shipped_table_parses():
    assert os.path.exists(TABLE)
    rc, n, _ = _load(TABLE)
    assert rc == N.MMI_OK and n > 0
    rc2, n2, _ = _load(None)  # the default: the table beside the library
    assert rc2 == N.MMI_OK and n2 == n


def test_rows_and_comments(tmp_path):
    p = tmp_path / "t.tsv"
    p.write_text("# comment\n\naffine f32 3 image 64 24 4 16 2  # trailing note\nbspline bf16 3 image 64 21 2 16 1\n")
    rc, n, _ = _load(str(p))
    assert rc == N.MMI_OK and n == 2


@pytest.mark.parametrize("line,what", [
    ("affine f32 3 image 64 24 4 16", "expected 9 fields"),
    ("rigid f32 3 image 64 24 4 16 2", "out of range"),
    ("affine f64 3 image 64 24 4 16 2", "out of range"),
    ("affine f32 4 image 64 24 4 16 2", "out of range"),
    ("affine f32 3 image 64 24 4 16 3", "out of range"),
    ("affine f32 3 image 64 24 0 16 2", "out of range"),
])
def test_bad_rows_are_refused(tmp_path, line, what):
    p = tmp_path / "bad.tsv"
    p.write_text("# ok\n" + line + "\n")
    rc, _, msg = _load(str(p))
    assert rc == N.MMI_ERR_ARGUMENT
    assert what in msg and ":2:" in msg


def test_missing_table(tmp_path):
    rc, _, msg = _load(str(tmp_path / "none.tsv"))
    assert rc == N.MMI_ERR_ARGUMENT and "cannot open" in msg
    assert _load(TABLE)[0] == N.MMI_OK  # restore the shipped table for later tests in this process
