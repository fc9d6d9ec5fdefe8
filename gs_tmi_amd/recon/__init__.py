"""Reconstruction harness around the HIP tile voxelizer.

The 3DGSMR loop (complex Gaussians -> volume -> coil maps -> centred 3D FFT ->
undersampling mask -> k-space loss), rebuilt in PyTorch on top of
``gs_tmi_amd.TileVoxelizer``: synthetic phantoms and coils, sampling masks,
metrics, the Gaussian parameter set with densification and pruning, and the
Adam training loop with time-to-NMSE tracking.  FFTs run through torch.fft
(hipFFT on ROCm).
"""

from .data import (MRIProblem, ellipsoid_phantom, expand_phase_mask, fft3c, gaussian_coils, ifft3c,  # noqa: F401
                   make_problem, poisson_disc_mask, variable_density_mask)
from .metrics import nmse, psnr  # noqa: F401
from .model import GaussianSet  # noqa: F401
from .trainer import ReconConfig, Reconstructor  # noqa: F401
