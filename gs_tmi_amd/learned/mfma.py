"""ctypes binding of libgstv_mlp.so (include/gstv_mlp.h) and a dense layer whose
three products run on the BF16 MFMA GEMM:

  forward   Y  = X W^T + b          gemm_nt(X, W)
  backward  dX = dY W               gemm_nt(dY, W^T)
            dW = dY^T X             gemm_nt(dY^T, X^T)

Operands are rounded to bf16 (round-to-nearest-even), products accumulate in
fp32, master weights and optimizer state stay fp32."""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch
import torch.nn as nn

LIB_PATH = Path(__file__).resolve().parents[1] / "_lib" / "libgstv_mlp.so"
ACT = {"none": 0, "relu": 1, "gelu": 2}

_lock = threading.Lock()
_lib = None

_P, _I, _I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
PROTOTYPES = {
    "gstv_mlp_status_string": (ctypes.c_char_p, [_I]),
    "gstv_mlp_gemm_nt_bf16": (_I, [_P, _I64, _P, _I64, _P, _I, _I, _I, _I, _P, _I64, _P, _I64, _P]),
    "gstv_mlp_transpose_bf16": (_I, [_P, _I, _I, _P, _P]),
    "gstv_mlp_cast_bf16": (_I, [_P, _I64, _P, _P]),
}


def load_library(path=None):
    lib = ctypes.CDLL(str(path or LIB_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


def get_lib():
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(f"{LIB_PATH} is not built (run `python -m gs_tmi_amd.build`)")
            _lib = load_library()
        return _lib


def _check(st, what):
    if st != 0:
        raise RuntimeError(f"{what} failed: {get_lib().gstv_mlp_status_string(st).decode()} (status {st})")


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


def to_bf16(x: torch.Tensor) -> torch.Tensor:
    """fp32 -> bf16 (RNE) on the GPU with the library's cast kernel."""
    x = x.contiguous().float()
    if not x.is_cuda:
        raise RuntimeError("MFMA layers need ROCm GPU tensors")
    out = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    with torch.cuda.device(x.device):
        _check(get_lib().gstv_mlp_cast_bf16(x.data_ptr(), x.numel(), out.data_ptr(), _stream(x)), "cast_bf16")
    return out


def transpose_bf16(x: torch.Tensor) -> torch.Tensor:
    x = x.contiguous()
    rows, cols = x.shape
    out = torch.empty(cols, rows, dtype=torch.bfloat16, device=x.device)
    with torch.cuda.device(x.device):
        _check(get_lib().gstv_mlp_transpose_bf16(x.data_ptr(), rows, cols, out.data_ptr(), _stream(x)),
               "transpose_bf16")
    return out


def gemm_nt(a: torch.Tensor, b: torch.Tensor, bias: torch.Tensor | None = None, act: str = "none",
            want_bf16: bool = False):
    """fp32 C = act(a @ b^T + bias) for bf16 a [M,K], b [N,K]; returns C (and its bf16 copy)."""
    assert a.dtype == torch.bfloat16 and b.dtype == torch.bfloat16 and a.is_contiguous() and b.is_contiguous()
    m, k = a.shape
    n, k2 = b.shape
    assert k == k2
    c = torch.empty(m, n, dtype=torch.float32, device=a.device)
    cb = torch.empty(m, n, dtype=torch.bfloat16, device=a.device) if want_bf16 else None
    bias_p = None if bias is None else bias.contiguous().float().data_ptr()
    with torch.cuda.device(a.device):
        st = get_lib().gstv_mlp_gemm_nt_bf16(a.data_ptr(), k, b.data_ptr(), k, bias_p, ACT[act], m, n, k,
                                             c.data_ptr(), n, None if cb is None else cb.data_ptr(), n, _stream(a))
    _check(st, "gemm_nt_bf16")
    return (c, cb) if want_bf16 else c


class _MfmaLinearFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, weight, bias):
        xb = x.contiguous() if x.dtype == torch.bfloat16 else to_bf16(x)
        wb = to_bf16(weight)
        y = gemm_nt(xb, wb, bias)
        ctx.save_for_backward(xb, wb)
        ctx.has_bias = bias is not None
        ctx.x_dtype = x.dtype
        return y

    @staticmethod
    def backward(ctx, gy):
        xb, wb = ctx.saved_tensors
        gyb = to_bf16(gy)
        gx = gw = gb = None
        if ctx.needs_input_grad[0]:
            gx = gemm_nt(gyb, transpose_bf16(wb))                    # [M, K]
            if ctx.x_dtype != gx.dtype:
                gx = gx.to(ctx.x_dtype)
        if ctx.needs_input_grad[1]:
            gw = gemm_nt(transpose_bf16(gyb), transpose_bf16(xb))    # [N, K]
        if ctx.has_bias and ctx.needs_input_grad[2]:
            gb = gy.sum(dim=0)
        return gx, gw, gb


class MfmaLinear(nn.Module):
    """Dense layer y = x W^T + b with BF16 MFMA products and fp32 master weights."""

    def __init__(self, in_features, out_features, bias=True, generator=None):
        super().__init__()
        bound = 1.0 / in_features ** 0.5
        w = (torch.rand(out_features, in_features, generator=generator) * 2 - 1) * bound
        self.weight = nn.Parameter(w)
        self.bias = nn.Parameter((torch.rand(out_features, generator=generator) * 2 - 1) * bound) if bias else None

    def forward(self, x):
        return _MfmaLinearFn.apply(x, self.weight, self.bias)
