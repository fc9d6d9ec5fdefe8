"""Magnitude PSNR and NMSE with the reference's definitions
(metrics/metrics.py:30-40 and :70-77)."""

from __future__ import annotations

import math

import torch


def _mag(x: torch.Tensor) -> torch.Tensor:
    return x.abs().double() if torch.is_complex(x) else x.double().abs()


def psnr(pred: torch.Tensor, target: torch.Tensor) -> float:
    """10 log10(range^2 / MSE) on magnitudes; range = max - min of the target magnitude."""
    p, t = _mag(pred), _mag(target)
    mse = float(torch.mean((p - t) ** 2))
    if mse == 0.0:
        return float("inf")
    rng = float(t.max() - t.min())
    if rng <= 0:
        rng = 1.0
    return 10.0 * math.log10(rng * rng / mse)


def nmse(pred: torch.Tensor, target: torch.Tensor) -> float:
    """sum (|pred| - |target|)^2 / sum |target|^2."""
    p, t = _mag(pred), _mag(target)
    den = float(torch.sum(t * t))
    return float(torch.sum((p - t) ** 2)) / den if den > 0 else float("inf")
