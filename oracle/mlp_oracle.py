"""FP64 NumPy oracle for the learned initialiser's network.  TEST INFRASTRUCTURE ONLY.

Only the tests may import this module, and only as the checker: the product
(``gs_tmi_amd.learned``) never calls it.

The initialiser has no counterpart in the reference (3DGSMR initialises from
the zero-filled image only, gaussian/gaussian_model.py:308-355), so this is a
restatement of the network the package defines, not of reference code:

* ``bf16_round``     fp32 -> bf16 round-to-nearest-even (ties to even), NaN kept.
* ``mlp_forward``    ReLU MLP, y = x W^T + b per layer, optionally with every
                     product operand rounded to bf16 the way the GPU layers
                     round them (activations and weights in the forward).
* ``mlp_backward``   the adjoint of ``mlp_forward``; with rounding it rounds the
                     operands of dX = dY W (dY, W) and dW = dY^T X (dY, X) as
                     the GPU backward does; bias gradient = column sums of dY.
* ``slot_loss``      existence BCE-with-logits (mean over slots) plus smooth-L1
                     (mean over the 8 fields of occupied slots), and its gradient.

Pinning (tests/test_mlp_oracle.py): ``bf16_round`` against torch's
``to(torch.bfloat16)`` and hand-worked ties; ``mlp_forward`` against explicit
loops; ``mlp_backward`` and ``slot_loss`` gradients against FP64 central finite
differences; ``slot_loss`` value against torch's BCE and smooth-L1 routines.
"""

from __future__ import annotations

import numpy as np


def bf16_round(x) -> np.ndarray:
    """Rounds to the nearest bf16 (ties to even) through the fp32 bit pattern;
    returns float64 holding bf16 values."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    is_nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(is_nan, (u | 0x00400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _rnd(x, on):
    return bf16_round(x) if on else np.asarray(x, dtype=np.float64)


def mlp_forward(params, x, round_bf16=False):
    """params: list of (W [out, in], b [out]); ReLU between layers, none after
    the last.  Returns (output, activations) where activations[i] is the input
    of layer i and masks[i] is the ReLU mask after layer i (i < last)."""
    acts, masks = [], []
    h = np.asarray(x, dtype=np.float64)
    for i, (w, b) in enumerate(params):
        acts.append(h)
        z = _rnd(h, round_bf16) @ _rnd(w, round_bf16).T + np.asarray(b, dtype=np.float64)
        if round_bf16:
            z = z.astype(np.float32).astype(np.float64)
        if i < len(params) - 1:
            masks.append(z > 0)
            h = np.maximum(z, 0.0)
        else:
            h = z
    return h, (acts, masks)


def mlp_backward(params, cache, gy, round_bf16=False):
    """Gradients [(dW, db), ...] and dX for upstream gradient gy of the output."""
    acts, masks = cache
    grads = [None] * len(params)
    g = np.asarray(gy, dtype=np.float64)
    for i in range(len(params) - 1, -1, -1):
        w, _ = params[i]
        gr = _rnd(g, round_bf16)
        dw = gr.T @ _rnd(acts[i], round_bf16)
        db = g.sum(axis=0)
        grads[i] = (dw, db)
        gx = gr @ _rnd(w, round_bf16)
        if i > 0:
            gx = gx * masks[i - 1]
        g = gx
    return grads, g


def slot_loss(pred, tgt, beta=0.1, w_exist=1.0):
    """pred, tgt: [..., 9] with field 8 the existence logit / label.  Returns
    (loss, dloss/dpred)."""
    pred = np.asarray(pred, dtype=np.float64)
    tgt = np.asarray(tgt, dtype=np.float64)
    z, e = pred[..., 8], tgt[..., 8]
    n = z.size
    # BCE with logits, stable form: max(z,0) - z e + log(1 + exp(-|z|)).
    bce = np.sum(np.maximum(z, 0) - z * e + np.log1p(np.exp(-np.abs(z)))) / n
    grad = np.zeros_like(pred)
    grad[..., 8] = w_exist * (1.0 / (1.0 + np.exp(-z)) - e) / n
    occ = e > 0.5
    m = int(occ.sum()) * 8
    reg = 0.0
    if m:
        d = pred[..., :8][occ] - tgt[..., :8][occ]
        ad = np.abs(d)
        reg = np.sum(np.where(ad < beta, 0.5 * d * d / beta, ad - 0.5 * beta)) / m
        gsub = np.where(ad < beta, d / beta, np.sign(d)) / m
        gfull = grad[..., :8]
        gfull[occ] = gsub
        grad[..., :8] = gfull
    return w_exist * bce + reg, grad


def central_fd(f, x, h=1e-6):
    """Central differences of scalar f at every entry of array x (modified in place, restored)."""
    g = np.zeros_like(x, dtype=np.float64)
    it = np.nditer(x, flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        old = x[i]
        x[i] = old + h
        fp = f()
        x[i] = old - h
        fm = f()
        x[i] = old
        g[i] = (fp - fm) / (2 * h)
    return g
