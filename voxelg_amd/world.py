"""Host-side mirror of voxelG's World and Camera (src/voxel.rs:328-379, src/camera.rs:4-42).

A World holds the window in the reference's own layout (72-byte bricks, x-fastest linear
indices, one u64 per tile / chunk / L4 cell, uniform-material byte tables), which is what
`DeviceWorld.upload` consumes, like Renderer::upload_world (src/renderer.rs:1248).
Generation and pyramid building run in libvoxelg_amd.so's C++ host code.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._lib import BRICK_DTYPE, check, f32x3, i32x3, lib, ptr

MAT_AIR = 0
MAT_SAND = 1
MAT_GRASS = 2
MAT_DIRT = 3
MAT_STONE = 4
MAT_WATER = 12
MAT_GLASS = 18
MAT_LEAF_FRINGE = 33
MAT_TURF = 35


def world_counts(dims) -> tuple[int, int, int, int, int]:
    """(bricks, tiles, chunks, l4 cells, voxels) of a window (src/world_dims.rs:9-37)."""
    out = (C.c_int64 * 5)()
    check(lib().vg_world_counts(i32x3(dims), out), "vg_world_counts")
    return tuple(int(x) for x in out)


def terrain_height(dims, seed: int, x: int, z: int) -> int:
    """Surface height of world column (x, z) of the benchmark terrain."""
    h = C.c_int32()
    check(lib().vg_terrain_height(i32x3(dims), C.c_uint64(seed), int(x), int(z), C.byref(h)),
          "vg_terrain_height")
    return int(h.value)


class World:
    """The voxel window: bricks + occupancy pyramid + uniform LUTs, reference layout."""

    def __init__(self, dims=(512, 256, 512), origin=(0, 0, 0)):
        self.dims = tuple(int(d) for d in dims)
        self.origin = tuple(int(o) for o in origin)
        nb, nt, nc, nl4, _ = world_counts(self.dims)
        self.bricks = np.zeros(nb, BRICK_DTYPE)
        self.tile_mask = np.zeros(nt, np.uint64)
        self.chunk_mask = np.zeros(nc, np.uint64)
        self.l4_mask = np.zeros(nl4, np.uint64)
        self.brick_uniform = np.zeros(nb, np.uint8)
        self.tile_uniform = np.zeros(nt, np.uint8)
        self.brick_dims = tuple(d // 4 for d in self.dims)

    # ---------------------------------------------------------------- build --
    def generate(self, seed: int, threads: int = 0) -> "World":
        """Benchmark terrain (fBm heightmap, caves, grass/dirt/stone/sand, water < y 56)."""
        L = lib()
        check(L.vg_generate_terrain(i32x3(self.dims), i32x3(self.origin), C.c_uint64(seed),
                                    ptr(self.bricks), threads), "vg_generate_terrain")
        return self.build_pyramid(threads)

    def build_pyramid(self, threads: int = 0) -> "World":
        """rebuild_all_masks + rebuild_all_uniform (src/voxel.rs:950-1000)."""
        check(lib().vg_build_pyramid(i32x3(self.dims), ptr(self.bricks), ptr(self.tile_mask),
                                     ptr(self.chunk_mask), ptr(self.l4_mask),
                                     ptr(self.brick_uniform), ptr(self.tile_uniform), threads),
              "vg_build_pyramid")
        return self

    def terrain_height(self, seed: int, x: int, z: int) -> int:
        return terrain_height(self.dims, seed, x, z)

    # ---------------------------------------------------------------- edits --
    def _slots(self, xyz):
        xyz = np.asarray(xyz, np.int64).reshape(-1, 3)
        o = np.asarray(self.origin, np.int64)
        d = np.asarray(self.dims, np.int64)
        rel = xyz - o
        inside = np.all((rel >= 0) & (rel < d), axis=1)
        s = np.mod(xyz, d)
        bx, by, bz = s[:, 0] >> 2, s[:, 1] >> 2, s[:, 2] >> 2
        BX, BY, _ = self.brick_dims
        bi = bx + by * BX + bz * BX * BY
        vi = (s[:, 0] & 3) + (s[:, 2] & 3) * 4 + (s[:, 1] & 3) * 16
        return inside, bi, vi

    def set_voxels(self, xyz, mats, rebuild: bool = True) -> "World":
        """apply_edit for many WORLD voxels (src/voxel.rs:450-474, Brick::set :311-324):
        mat 0 clears. Voxels outside the window are ignored. Rebuilds the pyramid."""
        inside, bi, vi = self._slots(xyz)
        mats = np.broadcast_to(np.asarray(mats, np.uint8), inside.shape)
        bi, vi, mats = bi[inside], vi[inside], mats[inside]
        self.bricks["materials"][bi, vi] = mats
        touched = np.unique(bi)
        occ = (self.bricks["materials"][touched] != 0).astype(np.uint64)
        self.bricks["occupancy"][touched] = (occ << np.arange(64, dtype=np.uint64)).sum(
            axis=1, dtype=np.uint64)
        if rebuild:
            self.build_pyramid()
        return self

    def material_at(self, xyz) -> np.ndarray:
        """World::material_at_world (src/voxel.rs:1136-1156) for many voxels."""
        inside, bi, vi = self._slots(xyz)
        out = np.zeros(inside.shape, np.uint8)
        m = self.bricks["materials"][bi[inside], vi[inside]]
        occ = (self.bricks["occupancy"][bi[inside]] >> vi[inside].astype(np.uint64)) & np.uint64(1)
        out[inside] = np.where(occ == 1, m, 0)
        return out

    def top_surface(self) -> np.ndarray:
        """Window-local y of the highest solid voxel of every (x, z) column (-1 if none),
        shape (X, Z)."""
        X, Y, Z = self.dims
        BX, BY, BZ = self.brick_dims
        occ = np.ascontiguousarray(self.bricks["occupancy"]).reshape(BZ, BY, BX)
        bits = np.unpackbits(occ.view(np.uint8).reshape(BZ, BY, BX, 8), axis=-1,
                             bitorder="little").reshape(BZ, BY, BX, 4, 4, 4)  # ly, lz, lx
        vox = bits.transpose(2, 5, 1, 3, 0, 4).reshape(X, Y, Z)  # x, y, z
        has = vox.any(axis=1)
        top = Y - 1 - np.argmax(vox[:, ::-1, :], axis=1)
        return np.where(has, top, -1)

    def solid_fraction(self) -> float:
        occ = np.ascontiguousarray(self.bricks["occupancy"])
        bits = np.unpackbits(occ.view(np.uint8)).sum(dtype=np.int64)
        return float(bits) / float(self.dims[0] * self.dims[1] * self.dims[2])


@dataclass
class Camera:
    """Fly camera (src/camera.rs:4-42); `uniform()` is CameraUniform::from_camera."""
    pos: tuple = (256.0, 80.0, 256.0)
    yaw: float = 0.0
    pitch: float = -0.5
    fov_y: float = math.radians(70.0)

    def uniform(self, width: int, height: int, world_origin=(0, 0, 0)) -> np.ndarray:
        """176-byte CameraUniform as 44 float32 words (world_origin as int32 at 20..22)."""
        out = np.zeros(44, np.float32)
        check(lib().vg_camera_from_pose(f32x3(self.pos), self.yaw, self.pitch, self.fov_y,
                                        int(width), int(height), i32x3(world_origin),
                                        ptr(out)), "vg_camera_from_pose")
        return out


def sun_direction(elevation_rad: float, azimuth_rad: float) -> np.ndarray:
    """Unit sun vector at an elevation above the horizon and an azimuth about +y."""
    ce = math.cos(elevation_rad)
    return np.array([ce * math.sin(azimuth_rad), math.sin(elevation_rad),
                     ce * math.cos(azimuth_rad)], np.float32)
