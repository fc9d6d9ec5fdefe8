"""Pins of the FP64 MLP oracle (CPU): bf16 rounding, forward, backward, slot loss."""

from __future__ import annotations

import numpy as np
import torch

from oracle import mlp_oracle as mo


def test_bf16_round_matches_torch_and_ties():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=10000) * 10.0 ** rng.integers(-30, 30, size=10000),
                        [0.0, -0.0, np.inf, -np.inf, 3.0e38, 1e-40, -1e-42]]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).float().double().numpy()
    got = mo.bf16_round(x)
    assert np.array_equal(got, want)
    # Ties go to the even neighbour: 1 + 2^-8 is halfway between 1 and 1 + 2^-7.
    assert mo.bf16_round(np.float32(1 + 2 ** -8)) == 1.0
    assert mo.bf16_round(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6
    assert mo.bf16_round(np.float32(1 + 2 ** -8 + 2 ** -20)) == 1 + 2 ** -7
    assert np.isnan(mo.bf16_round(np.float32(np.nan)))


def _params(dims, rng):
    return [(rng.normal(size=(b, a)) / np.sqrt(a), rng.normal(size=b) * 0.1) for a, b in zip(dims[:-1], dims[1:])]


def test_forward_matches_explicit_loops():
    rng = np.random.default_rng(1)
    params = _params([5, 4, 3], rng)
    x = rng.normal(size=(2, 5))
    y, _ = mo.mlp_forward(params, x)
    for n in range(2):
        h = x[n]
        for li, (w, b) in enumerate(params):
            z = [sum(h[k] * w[j, k] for k in range(w.shape[1])) + b[j] for j in range(w.shape[0])]
            h = np.array([max(v, 0.0) for v in z]) if li == 0 else np.array(z)
        assert np.allclose(y[n], h, rtol=1e-14, atol=1e-14)


def test_backward_matches_central_differences():
    rng = np.random.default_rng(2)
    params = _params([6, 5, 4, 3], rng)
    x = rng.normal(size=(4, 6))
    gy = rng.normal(size=(4, 3))

    def f():
        y, _ = mo.mlp_forward(params, x)
        return float(np.sum(y * gy))

    y, cache = mo.mlp_forward(params, x)
    grads, gx = mo.mlp_backward(params, cache, gy)
    assert np.allclose(gx, mo.central_fd(f, x), rtol=1e-7, atol=1e-8)
    for (w, b), (dw, db) in zip(params, grads):
        assert np.allclose(dw, mo.central_fd(f, w), rtol=1e-7, atol=1e-8)
        assert np.allclose(db, mo.central_fd(f, b), rtol=1e-7, atol=1e-8)


def test_rounded_backward_rounds_the_product_operands():
    # With rounding on and exactly representable data, rounding changes nothing.
    rng = np.random.default_rng(3)
    params = [(mo.bf16_round(w), b) for w, b in _params([8, 8, 4], rng)]
    x = mo.bf16_round(rng.normal(size=(3, 8)))
    gy = mo.bf16_round(rng.normal(size=(3, 4)))
    y0, c0 = mo.mlp_forward(params, x)
    y1, c1 = mo.mlp_forward(params, x, round_bf16=True)
    # Layer 1's input is a product, so it is no longer bf16; only layer 0 agrees exactly.
    assert np.allclose(c0[0][1], c1[0][1], rtol=1e-6)
    g0, _ = mo.mlp_backward(params, c0, gy)
    assert np.allclose(g0[-1][1], gy.sum(0))


def test_slot_loss_matches_torch_and_fd():
    rng = np.random.default_rng(4)
    pred = rng.normal(size=(3, 5, 9))
    tgt = rng.normal(size=(3, 5, 9))
    tgt[..., 8] = (rng.random(size=(3, 5)) > 0.5).astype(float)
    loss, grad = mo.slot_loss(pred, tgt)
    p, t = torch.from_numpy(pred), torch.from_numpy(tgt)
    occ = t[..., 8] > 0.5
    want = (torch.nn.functional.binary_cross_entropy_with_logits(p[..., 8], t[..., 8])
            + torch.nn.functional.smooth_l1_loss(p[..., :8][occ], t[..., :8][occ], beta=0.1))
    assert abs(loss - float(want)) < 1e-12
    fd = mo.central_fd(lambda: mo.slot_loss(pred, tgt)[0], pred, h=1e-7)
    assert np.allclose(grad, fd, rtol=1e-5, atol=1e-7)
