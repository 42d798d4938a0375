"""Benchmark workloads: seeded worlds, cameras and sun directions.

small  256x128x256 world (4x2x4 chunks of 64^3), 1280x720 primary rays, tmax 512, LOD off.
large  2048x256x2048 world (32x4x32 chunks), 1920x1080 primary rays (tmax 700 = MAX_RAY_DIST,
       shaders/raymarch.wgsl:4501) with brick/tile LOD past 400/520 and the fused beam
       prologue, plus one sun shadow ray
       per hit (tmax 480 = SHADOW_MAX_DIST, :4517; sun elevation 35 deg, seeded azimuth).
Cameras: seeded (x, z) over open ground, height 90, random yaw, pitch -0.35
(src/renderer.rs:3539), vertical fov 70 deg (src/camera.rs:19).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._lib import VG_RAY_BEAM, VG_RAY_LOD
from .world import Camera, World, sun_direction, terrain_height

MAX_RAY_DIST = 700.0
SHADOW_MAX_DIST = 480.0


class SplitMix64:
    """Deterministic 64-bit generator (Steele, Lea, Flood 2014)."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


@dataclass(frozen=True)
class FrameWorkload:
    name: str
    dims: tuple
    width: int
    height: int
    primary_tmax: float
    primary_flags: int
    shadow: bool
    shadow_tmax: float = SHADOW_MAX_DIST
    seed: int = 0x5EED_0001

    def world(self, threads: int = 0) -> World:
        return World(self.dims).generate(self.seed, threads)

    def camera(self) -> tuple[Camera, np.ndarray]:
        """Seeded camera over open ground (terrain below y = 86 under it)."""
        rng = SplitMix64(self.seed ^ 0xCA3E_7A11)
        for _ in range(4096):
            x = self.dims[0] * (0.25 + 0.5 * rng.uniform())
            z = self.dims[2] * (0.25 + 0.5 * rng.uniform())
            if terrain_height(self.dims, self.seed, int(x), int(z)) < 86:
                break
        else:
            raise RuntimeError("no open ground found for the camera")
        yaw = 2.0 * math.pi * rng.uniform()
        cam = Camera(pos=(float(np.float32(x)), 90.0, float(np.float32(z))), yaw=yaw,
                     pitch=-0.35, fov_y=math.radians(70.0))
        return cam, cam.uniform(self.width, self.height)

    def sun(self) -> np.ndarray:
        rng = SplitMix64(self.seed ^ 0x5A11_D1E)
        return sun_direction(math.radians(35.0), 2.0 * math.pi * rng.uniform())


SMALL = FrameWorkload("small", (256, 128, 256), 1280, 720, 512.0, 0, False)
LARGE = FrameWorkload("large", (2048, 256, 2048), 1920, 1080, MAX_RAY_DIST,
                      VG_RAY_LOD | VG_RAY_BEAM, True)
