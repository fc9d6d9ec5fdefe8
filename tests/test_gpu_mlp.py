"""BF16 MFMA GEMM kernels and the MFMA dense layers against the FP64 oracle (GPU).

Tolerances: products of bf16-rounded operands accumulate in fp32, so against
an FP64 product of the same rounded operands each element is within
2^-22 * K * sum_k |a_k b_k| (checked with K-scaled slack); whole-network
gradients against the unrounded FP64 network are within 2e-2 relative L2
(bf16 operand rounding, 2^-9 relative per operand, over four layers).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _mfma():
    from gs_tmi_amd.learned import mfma

    return mfma


def _bf16_np(t: torch.Tensor) -> np.ndarray:
    return t.float().cpu().double().numpy()


def _check_product(got, a64, b64, bias=None, act="none"):
    exact = a64 @ b64.T + (0 if bias is None else bias[None, :])
    bound = np.abs(a64) @ np.abs(b64).T * (a64.shape[1] + 8) * 2.0 ** -23 + 1e-30
    if act == "relu":
        exact = np.maximum(exact, 0)
    assert np.all(np.abs(got - exact) <= bound + 1e-6 * np.abs(exact)), float(np.max(np.abs(got - exact) - bound))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 32), (17, 33, 40), (128, 128, 32), (130, 257, 1000),
                                   (300, 64, 13), (512, 288, 512), (2048, 512, 1024)])
def test_gemm_nt_matches_fp64_of_rounded_operands(m, n, k):
    mf = _mfma()
    print(f"[gemm] m={m} n={n} k={k}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(m * 7 + n * 3 + k)
    a = torch.randn(m, k, generator=g).to(DEV)
    b = torch.randn(n, k, generator=g).to(DEV)
    ab, bb = mf.to_bf16(a), mf.to_bf16(b)
    c = mf.gemm_nt(ab, bb)
    torch.cuda.synchronize()
    _check_product(c.cpu().double().numpy(), _bf16_np(ab), _bf16_np(bb))


def test_gemm_nt_unaligned_strides_and_bias_activation():
    mf = _mfma()
    g = torch.Generator(device="cpu").manual_seed(11)
    # Leading dimensions that are not multiples of 8 take the element-wise load path.
    big_a = torch.randn(70, 45, generator=g).to(DEV)
    big_b = torch.randn(50, 45, generator=g).to(DEV)
    ab, bb = mf.to_bf16(big_a), mf.to_bf16(big_b)
    bias = torch.randn(50, generator=g).to(DEV)
    lib = mf.get_lib()
    for act, name in ((1, "relu"), (0, "none")):
        c = torch.empty(70, 50, device=DEV)
        cb = torch.empty(70, 50, dtype=torch.bfloat16, device=DEV)
        st = lib.gstv_mlp_gemm_nt_bf16(ab.data_ptr(), 45, bb.data_ptr(), 45, bias.data_ptr(), act, 70, 50, 45,
                                       c.data_ptr(), 50, cb.data_ptr(), 50, torch.cuda.current_stream().cuda_stream)
        assert st == 0
        torch.cuda.synchronize()
        _check_product(c.cpu().double().numpy(), _bf16_np(ab), _bf16_np(bb), bias.cpu().double().numpy(), name)
        # The bf16 output is the RNE rounding of the fp32 output, bit for bit.
        assert torch.equal(cb.cpu(), c.cpu().to(torch.bfloat16))


def test_gemm_nt_sub_matrix_view():
    mf = _mfma()
    g = torch.Generator(device="cpu").manual_seed(12)
    a = mf.to_bf16(torch.randn(40, 72, generator=g).to(DEV))
    b = mf.to_bf16(torch.randn(24, 72, generator=g).to(DEV))
    lib = mf.get_lib()
    # A is rows 3.. and columns 8..8+48 of a 40 x 72 matrix (lda 72), B likewise.
    c = torch.empty(37, 24, device=DEV)
    st = lib.gstv_mlp_gemm_nt_bf16(a.data_ptr() + (3 * 72 + 8) * 2, 72, b.data_ptr() + 8 * 2, 72, None, 0, 37, 24,
                                   48, c.data_ptr(), 24, None, 0, torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    _check_product(c.cpu().double().numpy(), _bf16_np(a)[3:, 8:56], _bf16_np(b)[:, 8:56])


def test_cast_and_transpose_exact():
    mf = _mfma()
    g = torch.Generator(device="cpu").manual_seed(13)
    x = torch.randn(333, 77, generator=g) * 10.0 ** torch.randint(-20, 20, (333, 77), generator=g)
    xb = mf.to_bf16(x.to(DEV))
    torch.cuda.synchronize()
    assert torch.equal(xb.cpu(), x.to(torch.bfloat16))
    assert np.array_equal(_bf16_np(xb), mo.bf16_round(x.numpy()))
    t = mf.transpose_bf16(xb)
    torch.cuda.synchronize()
    assert torch.equal(t.cpu(), xb.cpu().t().contiguous())


def test_mfma_linear_gradients_match_rounded_oracle():
    mf = _mfma()
    torch.manual_seed(0)
    layer = mf.MfmaLinear(96, 40).to(DEV)
    x = torch.randn(50, 96, device=DEV, requires_grad=True)
    gy = torch.randn(50, 40, device=DEV)
    y = layer(x)
    y.backward(gy)
    torch.cuda.synchronize()
    params = [(layer.weight.detach().cpu().double().numpy(), layer.bias.detach().cpu().double().numpy())]
    y64, cache = mo.mlp_forward(params, x.detach().cpu().double().numpy(), round_bf16=True)
    grads, gx64 = mo.mlp_backward(params, cache, gy.cpu().double().numpy(), round_bf16=True)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    assert rel(y.detach().cpu().double().numpy(), y64) < 1e-5
    assert rel(x.grad.cpu().double().numpy(), gx64) < 1e-5
    assert rel(layer.weight.grad.cpu().double().numpy(), grads[0][0]) < 1e-5
    assert rel(layer.bias.grad.cpu().double().numpy(), grads[0][1]) < 1e-6


def test_init_mlp_gradients_match_fp64_oracle():
    from gs_tmi_amd.learned import InitMLP

    model = InitMLP(seed=3).to(DEV)
    g = torch.Generator(device="cpu").manual_seed(5)
    x = torch.randn(64, 1024, generator=g).to(DEV)
    gy = torch.randn(64, 288, generator=g).to(DEV)
    out = model(x)
    out.reshape(64, 288).backward(gy)
    torch.cuda.synchronize()
    params = [(l.weight.detach().cpu().double().numpy(), l.bias.detach().cpu().double().numpy()) for l in model.layers]
    x64, gy64 = x.cpu().double().numpy(), gy.cpu().double().numpy()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    # Against the oracle that rounds the same operands: fp32-accumulation agreement
    # (rare bf16 rounding-boundary flips of intermediate values allowed for).
    y_r, cache_r = mo.mlp_forward(params, x64, round_bf16=True)
    grads_r, _ = mo.mlp_backward(params, cache_r, gy64, round_bf16=True)
    assert rel(out.detach().reshape(64, 288).cpu().double().numpy(), y_r) < 1e-4
    for l, (dw, db) in zip(model.layers, grads_r):
        assert rel(l.weight.grad.cpu().double().numpy(), dw) < 1e-3
        assert rel(l.bias.grad.cpu().double().numpy(), db) < 1e-3
    # Against the exact FP64 network: the output differs by bf16 operand rounding
    # only; the gradients also by the hidden units whose pre-activation is within
    # rounding of 0 (their ReLU mask flips), so they are held to direction
    # (cosine > 0.99) and 10% relative L2.
    y64, cache = mo.mlp_forward(params, x64)
    grads, _ = mo.mlp_backward(params, cache, gy64)
    assert rel(out.detach().reshape(64, 288).cpu().double().numpy(), y64) < 2e-2
    for l, (dw, db) in zip(model.layers, grads):
        gw = l.weight.grad.cpu().double().numpy()
        cos = float(np.sum(gw * dw) / (np.linalg.norm(gw) * np.linalg.norm(dw)))
        assert cos > 0.99 and rel(gw, dw) < 0.1
        assert rel(l.bias.grad.cpu().double().numpy(), db) < 0.1


def test_gemm_epilogue_mask_and_transposed_output():
    """gstv_mlp_gemm_nt_bf16_ex: the masked result is the plain result where
    the mask entry is a positive number and 0 elsewhere (negative, +-0, NaN
    entries), its bf16 copy is the RNE rounding of the fp32 result, and the
    transposed copy is exactly the transpose of the bf16 copy (ragged M for
    the 4-row write path, both tile variants)."""
    mf = _mfma()
    for m, n, k in ((70, 50, 45), (2048, 512, 512), (131, 288, 512)):
        print(f"[gemm ex] m={m} n={n} k={k}", flush=True)
        g = torch.Generator(device="cpu").manual_seed(m + n + k)
        a = mf.to_bf16(torch.randn(m, k, generator=g).to(DEV))
        b = mf.to_bf16(torch.randn(n, k, generator=g).to(DEV))
        bias = torch.randn(n, generator=g).to(DEV)
        mask_src = torch.randn(m, n, generator=g)
        mask_src[0, :4] = torch.tensor([0.0, -0.0, float("nan"), float("inf")])
        mask = mask_src.to(torch.bfloat16).to(DEV)
        plain = mf.gemm_nt(a, b, bias, act="relu")
        c, cb, ct = mf.gemm_nt(a, b, bias, act="relu", want_bf16=True, want_bf16_t=True, relu_mask=mask)
        torch.cuda.synchronize()
        keep = (mask.float() > 0).cpu()
        want = torch.where(keep, plain.cpu(), torch.zeros(()))
        assert torch.equal(c.cpu(), want)
        assert torch.equal(cb.cpu(), c.cpu().to(torch.bfloat16))
        assert torch.equal(ct.cpu(), cb.cpu().t().contiguous())
        only_t = mf.gemm_nt(a, b, bias, act="relu", want_fp32=False, want_bf16_t=True)
        torch.cuda.synchronize()
        assert torch.equal(only_t.cpu(), plain.cpu().to(torch.bfloat16).t().contiguous())


def test_cast_with_transpose_exact():
    mf = _mfma()
    for rows, cols in ((1, 1), (33, 70), (2048, 288)):
        x = torch.randn(rows, cols)
        x[0, 0] = float("nan")
        xb, xt = mf.cast_t(x.to(DEV))
        _, xt2 = mf.cast_t(x.to(DEV), plain=False)
        torch.cuda.synchronize()
        ref = x.to(torch.bfloat16)
        assert torch.equal(xb.cpu()[1:], ref[1:]) and torch.isnan(xb.cpu()[0, 0].float())
        assert torch.equal(xt.view(torch.int16).cpu(), xb.view(torch.int16).cpu().t().contiguous())
        assert torch.equal(xt2.view(torch.int16).cpu(), xt.view(torch.int16).cpu())


def test_relu_net_equals_composed_layers():
    """InitMLP's fused network (bias + ReLU in the GEMM epilogues, ReLU
    masks in the dX epilogues, transposed operands from the producers) gives
    the output and every gradient of MfmaLinear layers composed with
    torch.relu, bit for bit, with and without autograd."""
    from gs_tmi_amd.learned import InitMLP

    model = InitMLP(seed=4).to(DEV)
    g = torch.Generator(device="cpu").manual_seed(6)
    x = torch.randn(300, 1024, generator=g).to(DEV).requires_grad_()
    gy = torch.randn(300, 288, generator=g).to(DEV)
    out = model(x).reshape(300, 288)
    out.backward(gy)
    fused = [x.grad.clone()] + [t.grad.clone() for l in model.layers for t in (l.weight, l.bias)]
    for t in [x] + [t for l in model.layers for t in (l.weight, l.bias)]:
        t.grad = None
    h = x
    for i, layer in enumerate(model.layers):
        h = layer(h)
        if i < len(model.layers) - 1:
            h = torch.relu(h)
    h.backward(gy)
    composed = [x.grad] + [t.grad for l in model.layers for t in (l.weight, l.bias)]
    assert torch.equal(out.detach(), h.detach())
    for a, b in zip(fused, composed):
        assert torch.equal(a, b)
    with torch.no_grad():
        assert torch.equal(model(x).reshape(300, 288), out.detach())
        from gs_tmi_amd.learned.mfma import to_bf16
        xb = to_bf16(x)
        assert torch.equal(model(xb).reshape(300, 288), out.detach())


def test_gemm_nt_k_split_invariance():
    """A K-contiguous product equals the sum of the products of its two K halves
    (sub-matrix views with the full leading dimension), to fp32 accumulation."""
    mf = _mfma()
    g = torch.Generator(device="cpu").manual_seed(14)
    m, n, k, k1 = 300, 200, 768, 320
    a = mf.to_bf16(torch.randn(m, k, generator=g).to(DEV))
    b = mf.to_bf16(torch.randn(n, k, generator=g).to(DEV))
    lib = mf.get_lib()
    s = torch.cuda.current_stream().cuda_stream
    full = mf.gemm_nt(a, b)
    lo = torch.empty(m, n, device=DEV)
    hi = torch.empty(m, n, device=DEV)
    assert lib.gstv_mlp_gemm_nt_bf16(a.data_ptr(), k, b.data_ptr(), k, None, 0, m, n, k1, lo.data_ptr(), n, None, 0,
                                     s) == 0
    assert lib.gstv_mlp_gemm_nt_bf16(a.data_ptr() + 2 * k1, k, b.data_ptr() + 2 * k1, k, None, 0, m, n, k - k1,
                                     hi.data_ptr(), n, None, 0, s) == 0
    torch.cuda.synchronize()
    a64, b64 = _bf16_np(a), _bf16_np(b)
    bound = np.abs(a64) @ np.abs(b64).T * (k + 8) * 2.0 ** -23
    diff = np.abs(full.cpu().double().numpy() - (lo + hi).cpu().double().numpy())
    assert np.all(diff <= bound), float(np.max(diff - bound))


@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (65, 63, 130), (256, 128, 200), (2048, 288, 512),
                                   (512, 1024, 2048), (4096, 512, 1024), (8192, 520, 72), (4100, 4097, 96)])
def test_gemm_nt_tile_variants(m, n, k):
    """The block-tile variants against FP64 of the rounded operands, with bias,
    ReLU and the bf16 copy: 32x32 tiles with 128-wide k stages (below one
    64x64 block per CU; ragged K across a partial stage), 64x64 with 128-wide
    stages (4096 x 512), 128x128 (8192 x 520, ragged N) and 256x256 (the last
    case, ragged edges)."""
    mf = _mfma()
    print(f"[gemm variant] m={m} n={n} k={k}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(m + n + k)
    ab = mf.to_bf16(torch.randn(m, k, generator=g).to(DEV))
    bb = mf.to_bf16(torch.randn(n, k, generator=g).to(DEV))
    bias = torch.randn(n, generator=g).to(DEV)
    c, cb = mf.gemm_nt(ab, bb, bias, act="relu", want_bf16=True)
    torch.cuda.synchronize()
    _check_product(c.cpu().double().numpy(), _bf16_np(ab), _bf16_np(bb), bias.cpu().double().numpy(), "relu")
    assert torch.equal(cb.cpu(), c.cpu().to(torch.bfloat16))
