"""voxelg_amd — an MI355X-native voxel ray-query engine with voxelG's ray-query interface.

Host mirror of the reference types (World, Camera, PickHit) over libvoxelg_amd.so, whose
HIP kernels (gfx950) do all the tracing. See include/vg.h for the C ABI.
"""
from ._lib import (BRICK_DTYPE, HIT_DTYPE, IT_NAMES, RAY_DTYPE, VG_FACE_NONE, VG_RAY_ANY_HIT,
                   VG_RAY_BEAM, VG_RAY_LOD, VG_RAY_SKIP_TRANSPARENT, VGError, lib)
from .world import (MAT_AIR, MAT_DIRT, MAT_GLASS, MAT_GRASS, MAT_LEAF_FRINGE, MAT_SAND,
                    MAT_STONE, MAT_TURF, MAT_WATER, Camera, World, sun_direction, world_counts)

__all__ = [
    "BRICK_DTYPE", "HIT_DTYPE", "RAY_DTYPE", "IT_NAMES", "VG_FACE_NONE", "VG_RAY_ANY_HIT", "VG_RAY_BEAM",
    "VG_RAY_LOD", "VG_RAY_SKIP_TRANSPARENT", "VGError", "lib", "World", "Camera",
    "sun_direction", "world_counts", "MAT_AIR", "MAT_DIRT", "MAT_GLASS", "MAT_GRASS",
    "MAT_LEAF_FRINGE", "MAT_SAND", "MAT_STONE", "MAT_TURF", "MAT_WATER",
    "Context", "DeviceWorld", "PickHit", "raycast", "decode_hits",
]


def __getattr__(name):
    # device-side names import torch lazily
    if name in ("Context", "DeviceWorld", "PickHit", "raycast", "decode_hits"):
        from . import device
        return getattr(device, name)
    raise AttributeError(name)
