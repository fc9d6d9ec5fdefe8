"""Timing of the BF16 MFMA GEMM at the learned initialiser's own shapes.

    python -m gs_tmi_amd.learned.gemm_bench [--iters N]

Rows (M x N x K, C = A B^T with both operands K-contiguous):
  2048 x 512 x 1024   training forward, layer 1 (batch 2048)
  2048 x 512 x 512    training forward, layers 2-3
  2048 x 288 x 512    training forward, output layer
  512 x 1024 x 2048   training weight gradient, layer 1 (K = batch)
  512 x 512 x 2048    training weight gradient, layers 2-3
  16384 x 512 x 1024  inference over 16,384 patches, layer 1
and one whole training step of the network (forward + backward, batch 2048)
and one inference pass (16,384 patches).  Each row is timed with HIP events on
the launching stream over ``iters`` calls after warm-up, beside torch.matmul
on the same bf16 operands (hipBLASLt).  Prints one JSON object.
"""

from __future__ import annotations

import argparse
import json

import torch

ROWS = [(2048, 512, 1024), (2048, 512, 512), (2048, 288, 512), (512, 1024, 2048), (512, 512, 2048),
        (16384, 512, 1024)]
PEAK_BF16 = 2500.0  # TFLOP/s, MI355X dense bf16


def _time(fn, iters):
    for _ in range(3):
        fn()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(iters):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / iters  # ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=50)
    args = ap.parse_args()

    from . import mfma
    from .model import InitMLP

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(0)
    out = {"rows": []}
    for m, n, k in ROWS:
        a = mfma.to_bf16(torch.randn(m, k, generator=g).to(dev))
        b = mfma.to_bf16(torch.randn(n, k, generator=g).to(dev))
        c = torch.empty(m, n, dtype=torch.float32, device=dev)
        lib = mfma.get_lib()
        stream = torch.cuda.current_stream(dev).cuda_stream

        def ours():
            lib.gstv_mlp_gemm_nt_bf16_ex(a.data_ptr(), k, b.data_ptr(), k, None, 0, None, n, m, n, k, c.data_ptr(),
                                         n, None, n, None, m, stream)

        bt = b.t()

        def blas():
            torch.matmul(a, bt)

        t_ours, t_blas = _time(ours, args.iters), _time(blas, args.iters)
        flop = 2.0 * m * n * k
        row = {"shape": f"{m}x{n}x{k}", "ms": round(t_ours, 5), "tflops": round(flop / t_ours / 1e9, 1),
               "frac_of_bf16_peak": round(flop / t_ours / 1e9 / PEAK_BF16, 4),
               "hipblaslt_ms": round(t_blas, 5), "hipblaslt_tflops": round(flop / t_blas / 1e9, 1)}
        out["rows"].append(row)
        print(f"[gemm] {row}", flush=True)

    model = InitMLP(seed=0).to(dev)
    x = torch.randn(2048, 1024, generator=g).to(dev)
    gy = torch.randn(2048, 32, 9, generator=g).to(dev)

    def train_step():
        for p in model.parameters():
            p.grad = None
        model(x).backward(gy)

    xi = mfma.to_bf16(torch.randn(16384, 1024, generator=g).to(dev))

    def infer():
        with torch.no_grad():
            model(xi)

    dims = [1024, 512, 512, 512, 288]
    layer_flop = [2.0 * a * b for a, b in zip(dims[:-1], dims[1:])]  # per sample
    fwd_flop = sum(layer_flop)
    # Training: forward, every weight gradient, and the input gradients of
    # layers 2-4 (the patches need none).
    train_flop = 2 * fwd_flop + sum(layer_flop[1:])
    t_train = _time(train_step, args.iters)
    t_inf = _time(infer, args.iters)
    out["network"] = {
        "train_step_batch_2048_ms": round(t_train, 4),
        "train_step_tflops": round(train_flop * 2048 / t_train / 1e9, 1),
        "inference_16384_ms": round(t_inf, 4),
        "inference_tflops": round(fwd_flop * 16384 / t_inf / 1e9, 1),
    }
    print(f"[gemm] network {out['network']}", flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
