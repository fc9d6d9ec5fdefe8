"""ctypes binding of ``libgstv.so`` (C ABI declared in include/gstv.h).

torch is imported first on purpose: the library then binds to the HIP runtime
torch already loaded, so streams and device pointers are shared.  The library
is loaded from this package's ``_lib/`` directory, built in-tree by
``gs_tmi_amd.build``.  There is no fallback: if the library or a GPU is
missing, every entry point raises RuntimeError.
"""

from __future__ import annotations

import ctypes
import math
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libgstv.so"

_lock = threading.Lock()
_lib = None
_load_error: str | None = None

_c_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/gstv.h.
PROTOTYPES = {
    "gstv_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "gstv_version": (ctypes.c_int, []),
    "gstv_bin_count_workspace_bytes": (_sz, [_i32]),
    "gstv_bin_count": (ctypes.c_int, [_c_p, _c_p, _i32, _i32, _i32, _i32, _i32, _c_p, _c_p, _c_p, _sz, _c_p]),
    "gstv_bin_fill_workspace_bytes": (_sz, [_i32, _i64, _i32]),
    "gstv_bin_fill": (ctypes.c_int,
                      [_c_p, _c_p, _i32, _i32, _i32, _i32, _i32, _c_p, _i64, _c_p, _c_p, _c_p, _sz, _c_p]),
    "gstv_forward_workspace_bytes": (_sz, [_i32]),
    "gstv_forward": (ctypes.c_int,
                     [_c_p, _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _c_p, _i32, _i32, _i32, _i32, _i32,
                      _c_p, _c_p, _c_p, _sz, _c_p]),
    "gstv_backward_workspace_bytes": (_sz, [_i32]),
    "gstv_backward": (ctypes.c_int,
                      [_c_p, _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _c_p, _i32, _c_p, _c_p, _i32, _i32, _i32,
                       _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _sz, _c_p]),
    "gstv_backward_noalloc": (ctypes.c_int,
                              [_c_p, _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _c_p, _i32, _c_p, _c_p, _i32, _i32, _i32,
                               _i32, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "gstv_keep_workspace_bytes": (ctypes.c_size_t, [_i32]),
    "gstv_keep_prepare": (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _i32, _i32, _i32, _i32, _i32,
                                         _c_p, _c_p, _c_p, _sz, _c_p]),
    "gstv_keep_clear": (ctypes.c_int, [_c_p, _sz, _i32, _c_p]),
    "gstv_keep_prologue": (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i32, _i32, _i32, _c_p,
                                          _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _sz, _c_p]),
    "gstv_keep_bin_fill": (ctypes.c_int,
                           [_c_p, _sz, _c_p, _c_p, _i32, _i32, _i32, _i32, _i32, _c_p, ctypes.c_int64, _c_p, _c_p,
                            _c_p, _sz, _c_p]),
    "gstv_keep_bin_fill_capped": (ctypes.c_int,
                                  [_c_p, _sz, _c_p, _c_p, _i32, _i32, _i32, _i32, _i32, _c_p, _c_p, ctypes.c_int64,
                                   _c_p, _c_p, _c_p, _sz, _c_p]),
    "gstv_forward_keep": (ctypes.c_int,
                          [_c_p, _sz, _c_p, _i32, _c_p, _i32, _i32, _i32, _i32, _i32, _c_p, _c_p, _c_p, _c_p]),
    "gstv_backward_keep": (ctypes.c_int,
                           [_c_p, _sz, _c_p, _c_p, _c_p, _i32, _c_p, _i32, _c_p, _c_p, _i32, _i32, _i32, _i32,
                            _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _c_p]),
    "gstv_backward_keep_params": (ctypes.c_int,
                                  [_c_p, _sz, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _i32, _c_p, _c_p, _i32, _i32,
                                   _i32, _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _c_p]),
    "gstv_count_evaluations": (ctypes.c_int,
                               [_c_p, _c_p, _c_p, _c_p, _c_p, _i32, _c_p, _c_p, _i32, _i32, _i32, _i32, _i32,
                                _c_p, _c_p, _sz, _c_p]),
    "gstv_prologue": (ctypes.c_int, [_c_p, _c_p, _c_p, _i32, _i32, _i32, _i32, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "gstv_epilogue": (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i32, _c_p, _c_p, _c_p, _c_p]),
    "gstv_profile_begin": (ctypes.c_int, []),
    "gstv_profile_end": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]),
}


def load_library(path: Path | str | None = None):
    """Load the library and declare its prototypes (no GPU call is made)."""
    lib = ctypes.CDLL(str(path or LIB_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def get_lib(strict: bool = True):
    """Memoised loader with the reference's strict/non-strict contract
    (gaussian/tile_voxelizer.py:23-74): a failure is remembered; strict mode
    raises RuntimeError, non-strict returns None."""
    global _lib, _load_error
    with _lock:
        if _lib is not None:
            return _lib
        if _load_error is None:
            try:
                if not LIB_PATH.exists():
                    raise FileNotFoundError(f"{LIB_PATH} is not built (run `python -m gs_tmi_amd.build`)")
                _lib = load_library()
                return _lib
            except Exception as exc:  # noqa: BLE001 - any loader failure is reported the same way
                _load_error = f"Failed to load the HIP tile voxelizer library: {exc}"
        if strict:
            raise RuntimeError(_load_error)
        return None


def _reset_for_tests():
    global _lib, _load_error
    with _lock:
        _lib = None
        _load_error = None


def _check(status: int, what: str):
    if status != 0:
        msg = get_lib().gstv_status_string(status).decode()
        raise RuntimeError(f"{what} failed: {msg} (status {status})")


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def aligned16(t: torch.Tensor) -> torch.Tensor:
    """``t`` itself when its data starts on a 16-byte boundary, else a copy
    (the kernels read quaternions as 16-byte vectors)."""
    return t if t.data_ptr() % 16 == 0 else t.clone()


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def _require_gpu(t: torch.Tensor, name: str):
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a ROCm GPU tensor; the HIP tile voxelizer has no CPU path")


def num_tiles_for(shape, tile_size: int) -> int:
    d, h, w = shape
    return (-(-d // tile_size)) * (-(-h // tile_size)) * (-(-w // tile_size))


def bin_count(center_vox: torch.Tensor, radius_vox: torch.Tensor, shape, tile_size: int):
    """Phase one of binning.  Returns (gauss_offsets int32 [n+1], total int64 [1]) on the device."""
    lib = get_lib()
    _require_gpu(center_vox, "center_vox")
    dev = center_vox.device
    c = center_vox.detach().contiguous().float()
    r = radius_vox.detach().contiguous().float()
    n = c.shape[0]
    offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
    total = torch.empty(1, dtype=torch.int64, device=dev)
    ws_bytes = lib.gstv_bin_count_workspace_bytes(n)
    ws = _workspace(ws_bytes, dev)
    d, h, w = shape
    with torch.cuda.device(dev):
        st = lib.gstv_bin_count(_ptr(c), _ptr(r), n, d, h, w, tile_size, _ptr(offsets), _ptr(total), _ptr(ws),
                                ws_bytes, _stream(dev))
    _check(st, "gstv_bin_count")
    return offsets, total


def bin_fill(center_vox, radius_vox, shape, tile_size, offsets, total_pairs: int, kept=None, tile_ranges=None,
             capped: bool = False, pair_ids=None):
    """Phase two of binning.  Returns (tile_ranges int32 [T,2], pair_ids int32 [P]).
    ``kept``: a ``Kept`` whose ``keep_prepare`` counted the pairs (offsets
    from it): the pairs whose tile meets a Gaussian's extent box.
    ``tile_ranges``: an int32 [T, 2] array to fill (allocated here if None);
    ``pair_ids`` likewise (int32, ``total_pairs`` entries).
    ``capped`` (with ``kept``): ``total_pairs`` is a capacity and the kernels
    read the total from ``kept.total`` on the device (gstv_keep_bin_fill_capped):
    nothing waits for the host; ``pair_ids`` has ``total_pairs`` entries, and a
    total above the capacity leaves every tile list empty."""
    lib = get_lib()
    dev = center_vox.device
    c = center_vox.detach().contiguous().float()
    r = radius_vox.detach().contiguous().float()
    n = c.shape[0]
    num_tiles = num_tiles_for(shape, tile_size)
    if total_pairs >= 2**31 - 1:
        _check(3, "gstv_bin_fill")
    if capped and kept is None:
        raise ValueError("capped binning reads the total that keep_prepare / keep_prologue counted: pass kept")
    if tile_ranges is None:
        tile_ranges = torch.empty(num_tiles, 2, dtype=torch.int32, device=dev)
    if pair_ids is None:
        pair_ids = torch.empty(total_pairs, dtype=torch.int32, device=dev)
    elif pair_ids.dtype != torch.int32 or pair_ids.numel() != total_pairs or not pair_ids.is_contiguous():
        raise ValueError("pair_ids must be a contiguous int32 array of total_pairs entries")
    ws_bytes = lib.gstv_bin_fill_workspace_bytes(n, total_pairs, num_tiles)
    ws = _workspace(ws_bytes, dev)
    d, h, w = shape
    with torch.cuda.device(dev):
        if kept is None:
            st = lib.gstv_bin_fill(_ptr(c), _ptr(r), n, d, h, w, tile_size, _ptr(offsets), total_pairs,
                                   _ptr(tile_ranges), _ptr(pair_ids), _ptr(ws), ws_bytes, _stream(dev))
        elif capped:
            st = lib.gstv_keep_bin_fill_capped(_ptr(kept.ws), kept.ws.numel(), _ptr(c), _ptr(r), n, d, h, w,
                                               tile_size, _ptr(offsets), _ptr(kept.total), total_pairs,
                                               _ptr(tile_ranges), _ptr(pair_ids), _ptr(ws), ws_bytes, _stream(dev))
        else:
            st = lib.gstv_keep_bin_fill(_ptr(kept.ws), kept.ws.numel(), _ptr(c), _ptr(r), n, d, h, w, tile_size,
                                        _ptr(offsets), total_pairs, _ptr(tile_ranges), _ptr(pair_ids), _ptr(ws),
                                        ws_bytes, _stream(dev))
    name = "gstv_bin_fill" if kept is None else ("gstv_keep_bin_fill_capped" if capped else "gstv_keep_bin_fill")
    _check(st, name)
    return tile_ranges, pair_ids


def _rho_stride(rho_r, rho_i, keep):
    """Element stride shared by the 1-D density views rho_r / rho_i: 1 for
    planar arrays; with ``keep``, also 2 for the columns of an interleaved
    [n, 2] (complex64) array."""
    if rho_r.dim() != 1 or rho_i.dim() != 1:
        raise ValueError("rho_r and rho_i must be 1-D")
    if not keep and rho_r.numel() <= 1 and rho_i.numel() <= 1:
        # torch calls a one-element view contiguous whatever its stride (the
        # real part of a one-element complex tensor keeps stride 2); the one
        # address read does not depend on it.
        return 1
    st = rho_r.stride(0)
    if rho_i.stride(0) != st:
        raise ValueError("rho_r and rho_i must have the same stride")
    if st != 1 and not keep:
        raise ValueError("strided density views need the keep entry points")
    if st < 1:
        raise ValueError("density stride must be positive")
    return st


class PairCapacity:
    """Size of the pair arrays a voxelizer bins into without reading the pair
    total back first (``bin_fill(..., capped=True)``).

    Unknown until a first exact call reports a total; then ``total * headroom
    + slack`` pairs, scaled with the Gaussian count when that grows.  It grows
    when a total exceeds it (that call is redone with the exact size) and
    shrinks when totals fall below half of it."""

    def __init__(self, headroom: float = 1.25, slack: int = 1 << 16):
        self.headroom = headroom
        self.slack = slack
        self.cap = None
        self.n_ref = 0
        self.overflows = 0

    def capacity(self, n: int):
        """Capacity for a call on n Gaussians, or None (no total seen yet)."""
        if self.cap is None or n <= 0:
            return None
        scale = max(1.0, n / max(self.n_ref, 1))
        return min(int(math.ceil(self.cap * scale)), 2**31 - 2)

    def update(self, n: int, total: int, used=None):
        """Records the total of a call on n Gaussians (binned with capacity
        ``used``, or None for an exact call)."""
        if used is not None and total > used:
            self.overflows += 1
        want = int(math.ceil(total * self.headroom)) + self.slack
        if self.cap is None or (used is not None and total > used) or total > self.cap or want < self.cap // 2:
            self.cap = want
            self.n_ref = n


class Kept:
    """State the keep entry points hand from the forward to the adjoint of one
    call: the keep workspace (Gaussian records and moment accumulator) and the
    per-pair column words.  ``fresh``: the accumulator is zero (a second
    adjoint clears it again)."""

    __slots__ = ("ws", "n", "cols", "fresh", "stride", "offsets", "total")

    def __init__(self, ws, n, stride):
        self.ws = ws
        self.n = n
        self.cols = None
        self.fresh = False
        self.stride = stride
        self.offsets = None
        self.total = None


def keep_prepare(center_vox, cov_inv, rho_r, rho_i, half_shape, radius_vox=None, shape=None, tile_size=None):
    """Queues gstv_keep_prepare: the Gaussian records in a new keep workspace.
    rho_r / rho_i may be the columns of an interleaved [n, 2] fp32 array.
    With ``radius_vox``, ``shape`` and ``tile_size`` it also counts the pairs
    whose tile meets a Gaussian's extent box: ``kept.offsets`` (int32 [n+1])
    and ``kept.total`` (int64 [1]) on the device, for ``bin_fill(...,
    kept=kept)``.  Returns the ``Kept`` (its accumulator not yet cleared)."""
    lib = get_lib()
    _require_gpu(center_vox, "center_vox")
    dev = center_vox.device
    n = center_vox.shape[0]
    stride = _rho_stride(rho_r, rho_i, True)
    kept = Kept(_workspace(lib.gstv_keep_workspace_bytes(n), dev), n, stride)
    d = h = w = 0
    tile = 0
    if radius_vox is not None:
        d, h, w = shape
        tile = tile_size
        kept.offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
        kept.total = torch.empty(1, dtype=torch.int64, device=dev)
    with torch.cuda.device(dev):
        st = lib.gstv_keep_prepare(_ptr(center_vox), _ptr(radius_vox), _ptr(cov_inv), _ptr(rho_r), _ptr(rho_i),
                                   stride, _ptr(half_shape), n, d, h, w, tile, _ptr(kept.offsets),
                                   _ptr(kept.total), _ptr(kept.ws), kept.ws.numel(), _stream(dev))
    _check(st, "gstv_keep_prepare")
    return kept


def keep_prologue(positions, scales, rotations, rho_r, rho_i, shape, tile_size):
    """The fused call's first pass (gstv_keep_prologue): prologue, Gaussian
    records, the extent-box pair counts and a zeroed moment accumulator.
    rho_r / rho_i as in ``keep_prepare``.  Returns (center_vox, radius_vox,
    cov_inv, half_shape, kept), ``kept.offsets`` / ``kept.total`` set and
    ``kept.fresh`` true."""
    lib = get_lib()
    _require_gpu(positions, "positions")
    dev = positions.device
    p = positions.detach().contiguous().float()
    s = scales.detach().contiguous().float()
    q = aligned16(rotations.detach().contiguous().float())
    n = p.shape[0]
    stride = _rho_stride(rho_r, rho_i, True)
    center = torch.empty(n, 3, dtype=torch.float32, device=dev)
    radius = torch.empty(n, dtype=torch.float32, device=dev)
    cov = torch.empty(n, 3, 3, dtype=torch.float32, device=dev)
    hs = torch.empty(3, dtype=torch.float32, device=dev)
    kept = Kept(_workspace(lib.gstv_keep_workspace_bytes(n), dev), n, stride)
    kept.offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
    kept.total = torch.empty(1, dtype=torch.int64, device=dev)
    d, h, w = shape
    with torch.cuda.device(dev):
        st = lib.gstv_keep_prologue(_ptr(p), _ptr(s), _ptr(q), _ptr(rho_r), _ptr(rho_i), stride, n, d, h, w,
                                    tile_size, _ptr(center), _ptr(radius), _ptr(cov), _ptr(hs), _ptr(kept.offsets),
                                    _ptr(kept.total), _ptr(kept.ws), kept.ws.numel(), _stream(dev))
    _check(st, "gstv_keep_prologue")
    kept.fresh = True
    return center, radius, cov, hs, kept


def keep_clear(kept):
    """Queues gstv_keep_clear: zeroes ``kept``'s moment accumulator."""
    lib = get_lib()
    dev = kept.ws.device
    with torch.cuda.device(dev):
        st = lib.gstv_keep_clear(_ptr(kept.ws), kept.ws.numel(), kept.n, _stream(dev))
    _check(st, "gstv_keep_clear")
    kept.fresh = True


def forward(center_vox, cov_inv, rho_r, rho_i, tile_ranges, pair_ids, half_shape, tile_size, d, h, w,
            keep=False):
    """Planar forward.  Returns (vol_r, vol_i), each fp32 [d*h*w].  With
    ``keep`` (True, or a ``Kept`` from ``keep_prepare`` of the same inputs),
    returns (vol_r, vol_i, kept), where ``kept`` is what
    ``backward(..., kept=kept)`` of the same inputs reuses; rho_r / rho_i may
    then also be the two columns of one interleaved [n, 2] fp32 array
    (``torch.view_as_real`` of a complex64 density)."""
    lib = get_lib()
    _require_gpu(center_vox, "center_vox")
    dev = center_vox.device
    n = center_vox.shape[0]
    _rho_stride(rho_r, rho_i, keep is not False)
    vol_r = torch.empty(d * h * w, dtype=torch.float32, device=dev)
    vol_i = torch.empty(d * h * w, dtype=torch.float32, device=dev)
    if keep is not False:
        if keep is True:
            kept = keep_prepare(center_vox, cov_inv, rho_r, rho_i, half_shape)
            keep_clear(kept)
        else:
            kept = keep
        kept.cols = torch.empty(max(pair_ids.numel(), 1), dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            st = lib.gstv_forward_keep(_ptr(kept.ws), kept.ws.numel(), _ptr(tile_ranges), tile_ranges.shape[0],
                                       _ptr(pair_ids), n, tile_size, d, h, w, _ptr(vol_r), _ptr(vol_i),
                                       _ptr(kept.cols), _stream(dev))
        _check(st, "gstv_forward_keep")
        return vol_r, vol_i, kept
    ws_bytes = lib.gstv_forward_workspace_bytes(n)
    ws = _workspace(ws_bytes, dev)
    with torch.cuda.device(dev):
        st = lib.gstv_forward(_ptr(center_vox), _ptr(cov_inv), _ptr(rho_r), _ptr(rho_i), _ptr(tile_ranges),
                              tile_ranges.shape[0], _ptr(pair_ids), _ptr(half_shape), n, tile_size, d, h, w,
                              _ptr(vol_r), _ptr(vol_i), _ptr(ws), ws_bytes, _stream(dev))
    _check(st, "gstv_forward")
    return vol_r, vol_i


def backward(center_vox, cov_inv, rho_r, rho_i, tile_ranges, pair_ids, half_shape, grad_r, grad_i, tile_size,
             d, h, w, kept=None, grad_rho=None):
    """Adjoint.  Returns (grad_center [n,3], grad_cov_inv [n,9], grad_rho_r [n], grad_rho_i [n]).
    ``kept``: what ``forward(..., keep=...)`` of the same inputs returned.
    ``grad_rho`` (with ``kept`` only): an fp32 [n, 2] array that receives the
    density gradient interleaved (``torch.view_as_complex`` gives the complex64
    gradient); grad_rho_r / grad_rho_i are then its columns."""
    lib = get_lib()
    _require_gpu(center_vox, "center_vox")
    dev = center_vox.device
    n = center_vox.shape[0]
    g_c = torch.empty(n, 3, dtype=torch.float32, device=dev)
    g_m = torch.empty(n, 9, dtype=torch.float32, device=dev)
    stride = _rho_stride(rho_r, rho_i, kept is not None)
    if grad_rho is not None:
        if kept is None or stride != 2 or grad_rho.shape != (n, 2) or grad_rho.dtype != torch.float32 or \
                not grad_rho.is_contiguous():
            raise ValueError("grad_rho needs kept, an interleaved density and a contiguous fp32 [n, 2] array")
        g_rr, g_ri = grad_rho[:, 0], grad_rho[:, 1]
    else:
        if stride != 1:
            raise ValueError("an interleaved density needs grad_rho")
        g_rr = torch.empty(n, dtype=torch.float32, device=dev)
        g_ri = torch.empty(n, dtype=torch.float32, device=dev)
    if kept is not None:
        if kept.stride != stride:
            raise ValueError("density layout differs from the kept forward's")
        if not kept.fresh:
            keep_clear(kept)
        with torch.cuda.device(dev):
            st = lib.gstv_backward_keep(_ptr(kept.ws), kept.ws.numel(), _ptr(cov_inv), _ptr(half_shape),
                                        _ptr(tile_ranges), tile_ranges.shape[0], _ptr(pair_ids), n, _ptr(grad_r),
                                        _ptr(grad_i), tile_size, d, h, w, _ptr(g_c), _ptr(g_m), _ptr(g_rr),
                                        _ptr(g_ri), stride, _ptr(kept.cols), _stream(dev))
        kept.fresh = False
        _check(st, "gstv_backward_keep")
        return g_c, g_m, g_rr, g_ri
    ws_bytes = lib.gstv_backward_workspace_bytes(n)
    ws = _workspace(ws_bytes, dev)
    with torch.cuda.device(dev):
        st = lib.gstv_backward(_ptr(center_vox), _ptr(cov_inv), _ptr(rho_r), _ptr(rho_i), _ptr(tile_ranges),
                               tile_ranges.shape[0], _ptr(pair_ids), _ptr(half_shape), n, _ptr(grad_r),
                               _ptr(grad_i), tile_size, d, h, w, _ptr(g_c), _ptr(g_m), _ptr(g_rr), _ptr(g_ri),
                               _ptr(ws), ws_bytes, _stream(dev))
    _check(st, "gstv_backward")
    return g_c, g_m, g_rr, g_ri


NOALLOC_MAX_GAUSSIANS = 262144  # GSTV_NOALLOC_MAX_GAUSSIANS


def backward_noalloc(center_vox, cov_inv, rho_r, rho_i, tile_ranges, pair_ids, half_shape, grad_r, grad_i,
                     tile_size, d, h, w, out=None):
    """gstv_backward_noalloc: the adjoint without a workspace (the library's
    static device array; at most NOALLOC_MAX_GAUSSIANS Gaussians).  ``out``:
    optional preallocated (grad_center [n,3], grad_cov_inv [n,9], grad_rho_r [n],
    grad_rho_i [n]); returns them."""
    lib = get_lib()
    _require_gpu(center_vox, "center_vox")
    dev = center_vox.device
    n = center_vox.shape[0]
    if out is None:
        out = (torch.empty(n, 3, dtype=torch.float32, device=dev), torch.empty(n, 9, dtype=torch.float32, device=dev),
               torch.empty(n, dtype=torch.float32, device=dev), torch.empty(n, dtype=torch.float32, device=dev))
    g_c, g_m, g_rr, g_ri = out
    with torch.cuda.device(dev):
        st = lib.gstv_backward_noalloc(_ptr(center_vox), _ptr(cov_inv), _ptr(rho_r), _ptr(rho_i), _ptr(tile_ranges),
                                       tile_ranges.shape[0], _ptr(pair_ids), _ptr(half_shape), n, _ptr(grad_r),
                                       _ptr(grad_i), tile_size, d, h, w, _ptr(g_c), _ptr(g_m), _ptr(g_rr),
                                       _ptr(g_ri), _stream(dev))
    _check(st, "gstv_backward_noalloc")
    return g_c, g_m, g_rr, g_ri


def backward_params(center_vox, cov_inv, scales, rotations, tile_ranges, pair_ids, half_shape, grad_r, grad_i,
                    tile_size, d, h, w, kept, grad_rho):
    """Adjoint of the fused call down to the parameters (gstv_backward_keep_params).
    ``kept`` from ``forward(..., keep=...)``; ``grad_rho`` an fp32 [n, 2] array
    for the interleaved density gradient.  Returns (grad_positions [n,3],
    grad_scales [n,3], grad_rotations [n,4]); rotations must be 16-byte aligned."""
    lib = get_lib()
    _require_gpu(center_vox, "center_vox")
    dev = center_vox.device
    n = center_vox.shape[0]
    if kept.stride != 2 or grad_rho.shape != (n, 2) or grad_rho.dtype != torch.float32 or \
            not grad_rho.is_contiguous():
        raise ValueError("backward_params needs an interleaved density and a contiguous fp32 [n, 2] grad_rho")
    gp = torch.empty(n, 3, dtype=torch.float32, device=dev)
    gs = torch.empty(n, 3, dtype=torch.float32, device=dev)
    gq = torch.empty(n, 4, dtype=torch.float32, device=dev)
    if not kept.fresh:
        keep_clear(kept)
    g_rr, g_ri = grad_rho[:, 0], grad_rho[:, 1]
    with torch.cuda.device(dev):
        st = lib.gstv_backward_keep_params(_ptr(kept.ws), kept.ws.numel(), _ptr(cov_inv), _ptr(half_shape),
                                           _ptr(scales), _ptr(aligned16(rotations)), _ptr(tile_ranges),
                                           tile_ranges.shape[0],
                                           _ptr(pair_ids), n, _ptr(grad_r), _ptr(grad_i), tile_size, d, h, w,
                                           _ptr(gp), _ptr(gs), _ptr(gq), _ptr(g_rr), _ptr(g_ri), 2,
                                           _ptr(kept.cols), _stream(dev))
    kept.fresh = False
    _check(st, "gstv_backward_keep_params")
    return gp, gs, gq


def count_evaluations(center_vox, cov_inv, rho_r, rho_i, tile_ranges, pair_ids, half_shape, tile_size, d, h, w):
    """Exact (E_total, E_active, E_culled, E_adjoint) of the tile kernels for
    this binning (synchronises): the reference kernel's evaluations, those with
    q <= 9, and those the forward and the backward kernel perform."""
    lib = get_lib()
    dev = center_vox.device
    n = center_vox.shape[0]
    counts = torch.zeros(4, dtype=torch.int64, device=dev)
    ws_bytes = lib.gstv_forward_workspace_bytes(n)
    ws = _workspace(ws_bytes, dev)
    with torch.cuda.device(dev):
        st = lib.gstv_count_evaluations(_ptr(center_vox), _ptr(cov_inv), _ptr(rho_r), _ptr(rho_i),
                                        _ptr(tile_ranges), tile_ranges.shape[0], _ptr(pair_ids), _ptr(half_shape),
                                        n, tile_size, d, h, w, _ptr(counts), _ptr(ws), ws_bytes, _stream(dev))
    _check(st, "gstv_count_evaluations")
    return tuple(int(v) for v in counts.cpu().tolist())


def profile_begin():
    _check(get_lib().gstv_profile_begin(), "gstv_profile_begin")


def profile_end():
    """Returns {"forward": (ms_total, launches), "backward": (ms_total, launches)}."""
    fms, bms = ctypes.c_double(), ctypes.c_double()
    fl, bl = ctypes.c_int32(), ctypes.c_int32()
    _check(get_lib().gstv_profile_end(ctypes.byref(fms), ctypes.byref(fl), ctypes.byref(bms), ctypes.byref(bl)),
           "gstv_profile_end")
    return {"forward": (fms.value, fl.value), "backward": (bms.value, bl.value)}


def prologue(positions, scales, rotations, shape):
    """Fused prologue.  Returns (center_vox [n,3], radius_vox [n], cov_inv [n,3,3], half_shape [3])."""
    lib = get_lib()
    _require_gpu(positions, "positions")
    dev = positions.device
    p = positions.detach().contiguous().float()
    s = scales.detach().contiguous().float()
    q = aligned16(rotations.detach().contiguous().float())
    n = p.shape[0]
    center = torch.empty(n, 3, dtype=torch.float32, device=dev)
    radius = torch.empty(n, dtype=torch.float32, device=dev)
    cov = torch.empty(n, 3, 3, dtype=torch.float32, device=dev)
    hs = torch.empty(3, dtype=torch.float32, device=dev)
    d, h, w = shape
    with torch.cuda.device(dev):
        st = lib.gstv_prologue(_ptr(p), _ptr(s), _ptr(q), n, d, h, w, _ptr(center), _ptr(radius), _ptr(cov),
                               _ptr(hs), _stream(dev))
    _check(st, "gstv_prologue")
    return center, radius, cov, hs


def epilogue(scales, rotations, grad_center, grad_cov, shape):
    """Fused chain rule of the prologue.  Returns (grad_positions, grad_scales, grad_rotations)."""
    lib = get_lib()
    dev = scales.device
    n = scales.shape[0]
    gp = torch.empty(n, 3, dtype=torch.float32, device=dev)
    gs = torch.empty(n, 3, dtype=torch.float32, device=dev)
    gq = torch.empty(n, 4, dtype=torch.float32, device=dev)
    d, h, w = shape
    with torch.cuda.device(dev):
        st = lib.gstv_epilogue(_ptr(scales), _ptr(aligned16(rotations)), _ptr(grad_center.contiguous()),
                               _ptr(grad_cov.contiguous()), n, d, h, w, _ptr(gp), _ptr(gs), _ptr(gq), _stream(dev))
    _check(st, "gstv_epilogue")
    return gp, gs, gq
