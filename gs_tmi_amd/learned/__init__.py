"""Learned Gaussian initialiser: an MLP over 8^3 patches of the zero-filled
SENSE-adjoint image proposes an initial Gaussian set for the reconstruction.
Its dense layers run on hand-written BF16 MFMA GEMMs (FP32 accumulate); the
plain threshold initialisation stays the fallback."""

from .mfma import MfmaLinear, gemm_nt  # noqa: F401
from .model import InitMLP  # noqa: F401
