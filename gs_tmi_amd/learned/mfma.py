"""ctypes binding of libgstv_mlp.so (include/gstv_mlp.h), a dense layer whose
three products run on the BF16 MFMA GEMM, and a whole ReLU network on it:

  forward   Y  = X W^T + b          gemm_nt(X, W)
  backward  dX = dY W               gemm_nt(dY, W^T)
            dW = dY^T X             gemm_nt(dY^T, X^T)

Operands are rounded to bf16 (round-to-nearest-even), products accumulate in
fp32, master weights and optimizer state stay fp32.  The transposed operands
are written by the kernels that make the operands (the cast, or the GEMM
epilogue of the layer before), never by a separate transpose pass over
activations."""

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
    "gstv_mlp_gemm_nt_bf16_ex": (_I, [_P, _I64, _P, _I64, _P, _I, _P, _I64, _I, _I, _I, _P, _I64, _P, _I64, _P,
                                      _I64, _P]),
    "gstv_mlp_transpose_bf16": (_I, [_P, _I, _I, _P, _P]),
    "gstv_mlp_cast_bf16": (_I, [_P, _I64, _P, _P]),
    "gstv_mlp_cast_t_bf16": (_I, [_P, _I, _I, _P, _P, _P]),
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


def cast_t(x: torch.Tensor, plain: bool = True):
    """fp32 [R, C] -> (bf16 [R, C] or None, bf16 [C, R] transposed), one pass (RNE)."""
    x = x.contiguous().float()
    if not x.is_cuda or x.dim() != 2:
        raise RuntimeError("cast_t needs a 2-D ROCm GPU tensor")
    rows, cols = x.shape
    out = torch.empty(rows, cols, dtype=torch.bfloat16, device=x.device) if plain else None
    out_t = torch.empty(cols, rows, dtype=torch.bfloat16, device=x.device)
    with torch.cuda.device(x.device):
        _check(get_lib().gstv_mlp_cast_t_bf16(x.data_ptr(), rows, cols, None if out is None else out.data_ptr(),
                                              out_t.data_ptr(), _stream(x)), "cast_t_bf16")
    return out, out_t


def gemm_nt(a: torch.Tensor, b: torch.Tensor, bias: torch.Tensor | None = None, act: str = "none",
            want_bf16: bool = False, want_fp32: bool = True, want_bf16_t: bool = False,
            relu_mask: torch.Tensor | None = None):
    """C = act(a @ b^T + bias) for bf16 a [M,K], b [N,K], zeroed where
    ``relu_mask`` (bf16 [M,N]) is not positive.  Returns, in this order, those
    requested of: fp32 C, its bf16 copy, its bf16 transpose [N,M] (a single
    tensor when only one is requested)."""
    assert a.dtype == torch.bfloat16 and b.dtype == torch.bfloat16 and a.is_contiguous() and b.is_contiguous()
    m, k = a.shape
    n, k2 = b.shape
    assert k == k2
    dev = a.device
    c = torch.empty(m, n, dtype=torch.float32, device=dev) if want_fp32 else None
    cb = torch.empty(m, n, dtype=torch.bfloat16, device=dev) if want_bf16 else None
    ct = torch.empty(n, m, dtype=torch.bfloat16, device=dev) if want_bf16_t else None
    if c is None and cb is None and ct is None:
        raise ValueError("gemm_nt: no output requested")
    if relu_mask is not None:
        assert relu_mask.dtype == torch.bfloat16 and relu_mask.shape == (m, n) and relu_mask.is_contiguous()
    bias_p = None if bias is None else bias.contiguous().float().data_ptr()
    ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    with torch.cuda.device(dev):
        st = get_lib().gstv_mlp_gemm_nt_bf16_ex(a.data_ptr(), k, b.data_ptr(), k, bias_p, ACT[act], ptr(relu_mask),
                                                n, m, n, k, ptr(c), n, ptr(cb), n, ptr(ct), m, _stream(a))
    _check(st, "gemm_nt_bf16_ex")
    outs = [t for t in (c, cb, ct) if t is not None]
    return outs[0] if len(outs) == 1 else tuple(outs)


def _operand(x: torch.Tensor, want_t: bool):
    """bf16 operand of an fp32 or bf16 [R, C] input, and its transpose when
    ``want_t`` (one cast-and-transpose pass for fp32 input)."""
    if x.dtype == torch.bfloat16:
        xb = x.contiguous()
        return xb, (transpose_bf16(xb) if want_t else None)
    if want_t:
        return cast_t(x)
    return to_bf16(x), None


class _MfmaLinearFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, weight, bias):
        xb, xbt = _operand(x, ctx.needs_input_grad[1])
        wb, wbt = cast_t(weight)
        y = gemm_nt(xb, wb, bias)
        ctx.save_for_backward(xbt, wbt)
        ctx.has_bias = bias is not None
        ctx.x_dtype = x.dtype
        return y

    @staticmethod
    def backward(ctx, gy):
        xbt, wbt = ctx.saved_tensors
        gyb, gybt = cast_t(gy)
        gx = gw = gb = None
        if ctx.needs_input_grad[0]:
            gx = gemm_nt(gyb, wbt)      # [M, K]
            if ctx.x_dtype != gx.dtype:
                gx = gx.to(ctx.x_dtype)
        if ctx.needs_input_grad[1]:
            gw = gemm_nt(gybt, xbt)     # [N, K]
        if ctx.has_bias and ctx.needs_input_grad[2]:
            gb = gy.sum(dim=0)
        return gx, gw, gb


class _ReluNetFn(torch.autograd.Function):
    """A stack of dense layers with ReLU between them as one Function: every
    hidden layer is one GEMM whose epilogue adds the bias, applies the ReLU
    and writes the bf16 activation and its transpose (the next layer's input
    and this layer's weight-gradient operand); in backward each dX product
    applies the ReLU mask of the activation below in its epilogue and writes
    the gradient in fp32 (bias gradient), bf16 and transposed bf16.  Same
    arithmetic as composing ``MfmaLinear`` layers with ``torch.relu``.
    Arguments: x, then weight and bias of each layer."""

    @staticmethod
    def forward(ctx, x, *params):
        ws, bs = params[0::2], params[1::2]
        nl = len(ws)
        need_w = any(ctx.needs_input_grad[1:])
        h, ht = _operand(x, need_w and ctx.needs_input_grad[1])
        saved_h, saved_ht, saved_wt = [], [ht], []
        y = None
        for i, (w, b) in enumerate(zip(ws, bs)):
            wb, wbt = cast_t(w)
            saved_wt.append(wbt)
            if i < nl - 1:
                if need_w:
                    h, ht = gemm_nt(h, wb, b, act="relu", want_fp32=False, want_bf16=True, want_bf16_t=True)
                else:
                    h, ht = gemm_nt(h, wb, b, act="relu", want_fp32=False, want_bf16=True), None
                saved_h.append(h)
                saved_ht.append(ht)
            else:
                y = gemm_nt(h, wb, b)
        ctx.nl = nl
        ctx.x_dtype = x.dtype
        ctx.save_for_backward(*saved_h, *[t for t in saved_ht if t is not None], *saved_wt)
        ctx.has_ht = [t is not None for t in saved_ht]
        return y

    @staticmethod
    def backward(ctx, gy):
        nl = ctx.nl
        saved = list(ctx.saved_tensors)
        hs = saved[:nl - 1]
        rest = saved[nl - 1:]
        hts = []
        for has in ctx.has_ht:
            hts.append(rest.pop(0) if has else None)
        wts = rest
        grads = [None] * (2 * nl)
        g = gy.contiguous().float()
        gb16, gbt = cast_t(g)
        gx = None
        for i in range(nl - 1, -1, -1):
            if ctx.needs_input_grad[1 + 2 * i]:
                grads[2 * i] = gemm_nt(gbt, hts[i])
            if ctx.needs_input_grad[2 + 2 * i]:
                grads[2 * i + 1] = g.sum(dim=0)
            if i > 0:
                g, gb16, gbt = gemm_nt(gb16, wts[i], relu_mask=hs[i - 1], want_bf16=True, want_bf16_t=True)
            elif ctx.needs_input_grad[0]:
                gx = gemm_nt(gb16, wts[0])
                if ctx.x_dtype != gx.dtype:
                    gx = gx.to(ctx.x_dtype)
        return (gx, *grads)


def relu_net(x: torch.Tensor, layers) -> torch.Tensor:
    """``layers[-1](relu(... relu(layers[0](x))))`` for ``MfmaLinear`` layers
    with biases, through ``_ReluNetFn`` (without autograd: the forward GEMMs
    only, no transposed copies)."""
    params = []
    for layer in layers:
        if layer.bias is None:
            raise ValueError("relu_net needs layers with biases")
        params += [layer.weight, layer.bias]
    if torch.is_grad_enabled() and (x.requires_grad or any(p.requires_grad for p in params)):
        return _ReluNetFn.apply(x, *params)
    h = x.contiguous() if x.dtype == torch.bfloat16 else to_bf16(x)
    nl = len(layers)
    for i in range(nl):
        wb = to_bf16(params[2 * i])
        if i < nl - 1:
            h = gemm_nt(h, wb, params[2 * i + 1], act="relu", want_fp32=False, want_bf16=True)
        else:
            h = gemm_nt(h, wb, params[2 * i + 1])
    return h


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
