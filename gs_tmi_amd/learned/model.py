"""The initialiser network: 8^3 complex patch (1024 reals) -> 3 hidden layers
of 512 (ReLU) -> 32 slots x (3 offset + 3 log-scale + 2 amplitude + 1
existence logit) = 288 outputs."""

from __future__ import annotations

import torch
import torch.nn as nn

from .mfma import MfmaLinear, relu_net

PATCH = 8
SLOTS = 32
SLOT_FIELDS = 9  # offset (3), log-scale (3), amplitude (2), existence logit (1)


class InitMLP(nn.Module):
    def __init__(self, hidden=512, depth=3, seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        dims = [2 * PATCH ** 3] + [hidden] * depth + [SLOTS * SLOT_FIELDS]
        self.layers = nn.ModuleList(MfmaLinear(a, b, generator=g) for a, b in zip(dims[:-1], dims[1:]))

    def forward(self, x):
        # Bias and ReLU in the GEMM epilogues (relu_net); the same arithmetic
        # as the layers composed with torch.relu.
        x = relu_net(x, self.layers)
        return x.view(x.shape[0], SLOTS, SLOT_FIELDS)
