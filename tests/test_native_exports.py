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
    # first-pass keys and values (int32 per pair), digit histograms, per-block first Gaussians
    assert lib.gstv_bin_fill_workspace_bytes(10, 12345, 64) >= 2 * 4 * 12345
    # three or more radix passes (more than 2^16 tiles) ping-pong through a second pair
    assert lib.gstv_bin_fill_workspace_bytes(10, 12345, 1 << 17) >= 4 * 4 * 12345


MLP_HEADER = ROOT / "include" / "gstv_mlp.h"


def _declared_mlp():
    text = re.sub(r"/\*.*?\*/", "", MLP_HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(gstv_mlp_\w+)\s*\(", text)))


def test_mlp_library_exports_every_declared_symbol():
    from gs_tmi_amd.learned import mfma

    if not mfma.LIB_PATH.exists():
        pytest.skip("library not built (run `python -m gs_tmi_amd.build`)")
    lib = mfma.load_library()
    names = _declared_mlp()
    assert {"gstv_mlp_gemm_nt_bf16", "gstv_mlp_gemm_nt_bf16_ex", "gstv_mlp_transpose_bf16", "gstv_mlp_cast_bf16",
            "gstv_mlp_cast_t_bf16"} <= set(names)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(mfma.PROTOTYPES), "ctypes prototypes out of sync with the header"
    assert lib.gstv_mlp_status_string(0) == b"success"
    # Argument checks run on the host before any launch.
    assert lib.gstv_mlp_gemm_nt_bf16(None, 4, None, 4, None, 0, 2, 2, 4, None, 2, None, 2, None) == 1
    assert lib.gstv_mlp_gemm_nt_bf16(None, 2, None, 4, None, 0, 2, 2, 4, None, 2, None, 2, None) == 1
    assert lib.gstv_mlp_gemm_nt_bf16(None, 4, None, 4, None, 0, 0, 2, 4, 1, 2, None, 2, None) == 0


@pytest.mark.skipif(not _native.LIB_PATH.exists(), reason="library not built (run `python -m gs_tmi_amd.build`)")
def test_pair_count_limit_is_refused_before_any_allocation():
    """2^31 - 1 or more (tile, Gaussian) pairs cannot be indexed by the int32 pair ids:
    bin_fill raises the library's TOO_MANY_PAIRS message before allocating anything."""
    import torch

    with pytest.raises(RuntimeError, match="2\\^31"):
        _native.bin_fill(torch.zeros(1, 3), torch.zeros(1), (8, 8, 8), 8, torch.zeros(2, dtype=torch.int32), 2**31)


def test_density_stride_of_views():
    """Density views the entry points accept (host logic only, CPU tensors)."""
    import torch

    planar = torch.zeros(5)
    assert _native._rho_stride(planar, planar, keep=False) == 1
    inter = torch.view_as_real(torch.zeros(5, dtype=torch.complex64))
    assert _native._rho_stride(inter[:, 0], inter[:, 1], keep=True) == 2
    with pytest.raises(ValueError, match="strided"):
        _native._rho_stride(inter[:, 0], inter[:, 1], keep=False)
    # One element: torch calls the stride-2 real part of a one-element complex
    # tensor contiguous, and the one address read does not depend on the stride.
    one = torch.zeros(1, dtype=torch.complex64)
    re_part, im_part = one.real.contiguous(), one.imag.contiguous()
    assert re_part.stride(0) == 2
    assert _native._rho_stride(re_part, im_part, keep=False) == 1
    with pytest.raises(ValueError, match="1-D"):
        _native._rho_stride(torch.zeros(2, 2), torch.zeros(2, 2), keep=False)


def test_drop_in_differences_raise_clear_errors():
    """Where this package deliberately differs from the reference's
    TileVoxelizer (gaussian/tile_voxelizer.py:248-277), it raises instead of
    silently changing backend: use_cuda=False (the reference's PyTorch CPU
    path), a tile size below 1 (the reference would fail later, inside its
    binning), and -- on a machine without a ROCm GPU -- both strict and
    non-strict construction (the reference would fall back to its CPU path; a
    fallback here would not be the HIP path)."""
    import torch

    from gs_tmi_amd import TileVoxelizer

    with pytest.raises(RuntimeError, match="use_cuda=False"):
        TileVoxelizer((16, 16, 16), use_cuda=False)
    with pytest.raises(ValueError, match="tile_size must be a positive integer"):
        TileVoxelizer((16, 16, 16), tile_size=0)
    if not torch.cuda.is_available():
        for strict in (False, True):
            with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
                TileVoxelizer((16, 16, 16), strict_cuda=strict)
