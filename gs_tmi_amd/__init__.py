"""gs_tmi_amd: MI355X-native tile voxelizer for 3DGSMR (Qsdcm/GS_TMI).

Public interface (mirrors gaussian/tile_voxelizer.py of the reference):
``TileVoxelizer``, ``build_tile_pairs``, ``_TileVoxelizeHIP``.
"""

from .tile_voxelizer import (  # noqa: F401
    TileVoxelizer,
    _TileVoxelizeHIP,
    _load_hip_ext,
    _quaternion_to_rotation_matrix,
    build_tile_pairs,
    voxel_space_params,
)

__version__ = "0.1.0"
