"""The C-ABI library loads and exports every symbol include/gstv.h declares (CPU test).

Only host-side arithmetic entry points (workspace sizes, status strings) are
called here; nothing touches a GPU.
"""

from __future__ import annotations

import re

import pytest

from gs_tmi_amd import _native
from tests.conftest import ROOT

HEADER = ROOT / "include" / "gstv.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gstv_\w+)\s*\(", text)))


def test_header_declares_functions():
    names = _declared()
    assert "gstv_forward" in names and "gstv_backward" in names and "gstv_bin_fill" in names


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="library not built (run `python -m gs_tmi_amd.build`)")
def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared()) == set(_native.PROTOTYPES), "ctypes prototypes out of sync with the header"
    assert lib.gstv_version() >= 100
    assert lib.gstv_status_string(0) == b"success"
    assert b"2^31" in lib.gstv_status_string(3)
    assert lib.gstv_forward_workspace_bytes(1000) >= 64 * 1000
    assert lib.gstv_backward_workspace_bytes(1000) >= 64 * 1000 + 48 * 1000
    assert lib.gstv_bin_fill_workspace_bytes(10, 12345, 64) >= 4 * 4 * 12345
