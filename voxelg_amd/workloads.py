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

from ._lib import EDIT_DTYPE, RAY_DTYPE, VG_RAY_ANY_HIT, VG_RAY_BEAM, VG_RAY_LOD
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


@dataclass(frozen=True)
class BatchedWorkload:
    """Uneven batched ray queries on the small world: n_batches batches per step with sizes
    exp(U[ln lo, ln hi]); 40% coherent primary rays (64x64-pixel tiles of random cameras),
    30% shadow rays toward the sun and 30% cosine-hemisphere AO / bounce rays, both from
    random surface points; tmax uniform in {32, 128, 512, 1024}."""
    name: str = "batched"
    dims: tuple = (256, 128, 256)
    n_batches: int = 96
    size_lo: int = 256
    size_hi: int = 262144
    mix: tuple = (0.4, 0.3, 0.3)
    tmaxes: tuple = (32.0, 128.0, 512.0, 1024.0)
    seed: int = 0x5EED_0003

    def world(self, threads: int = 0) -> World:
        return World(self.dims).generate(SMALL.seed, threads)

    def rays(self, world: World):
        """(rays [total] RAY_DTYPE, batch sizes) for one step."""
        rng = np.random.default_rng(self.seed)
        sizes = np.exp(rng.uniform(np.log(self.size_lo), np.log(self.size_hi),
                                   self.n_batches)).astype(np.int64)
        total = int(sizes.sum())
        n_p = int(total * self.mix[0])
        n_s = int(total * self.mix[1])
        n_a = total - n_p - n_s
        prim = self._primary(rng, n_p)
        top = world.top_surface()
        shad = self._surface_rays(rng, world, top, n_s, "shadow")
        ao = self._surface_rays(rng, world, top, n_a, "ao")
        allr = np.concatenate([prim, shad, ao])
        # interleave the three kinds in chunks of 1024 rays, each chunk kept coherent
        chunks = np.arange(0, total, 1024)
        order = rng.permutation(len(chunks))
        perm = np.concatenate([np.arange(c, min(c + 1024, total)) for c in chunks[order]])
        allr = allr[perm]
        allr["tmax"] = rng.choice(np.asarray(self.tmaxes, np.float32), total)
        return allr, [int(x) for x in sizes]

    def _primary(self, rng, n):
        out = np.zeros(n, RAY_DTYPE)
        W, H, T = 1280, 720, 64
        done = 0
        while done < n:
            x = self.dims[0] * (0.25 + 0.5 * rng.random())
            z = self.dims[2] * (0.25 + 0.5 * rng.random())
            cam = Camera(pos=(x, 90.0, z), yaw=2 * math.pi * rng.random(), pitch=-0.35)
            u = cam.uniform(W, H)
            tx, ty = rng.integers(0, W // T), rng.integers(0, H // T)
            px, py = np.meshgrid(np.arange(tx * T, tx * T + T), np.arange(ty * T, ty * T + T))
            px, py = px.ravel().astype(np.float32), py.ravel().astype(np.float32)
            ndc_x = (px + 0.5) / W * 2 - 1
            ndc_y = 1 - (py + 0.5) / H * 2
            a = ndc_x * u[15] * (W / H)
            b = ndc_y * u[15]
            d = u[4:7][None] + u[8:11][None] * a[:, None] + u[12:15][None] * b[:, None]
            k = min(len(d), n - done)
            out["o"][done:done + k] = u[0:3]
            out["d"][done:done + k] = d[:k]
            done += k
        return out

    def _surface_rays(self, rng, world, top, n, kind):
        out = np.zeros(n, RAY_DTYPE)
        xs = rng.integers(0, self.dims[0], n)
        zs = rng.integers(0, self.dims[2], n)
        ys = top[xs, zs]
        ok = ys >= 0
        o = np.stack([xs + rng.random(n), ys + 1.0 + 1e-3, zs + rng.random(n)], 1)
        out["o"] = o.astype(np.float32)
        if kind == "shadow":
            out["d"] = LARGE.sun()
            out["flags"] = VG_RAY_ANY_HIT
        else:
            r1, phi = rng.random(n), 2 * math.pi * rng.random(n)
            s = np.sqrt(r1)
            out["d"] = np.stack([s * np.cos(phi), np.sqrt(1 - r1), s * np.sin(phi)], 1)
        out["d"][~ok] = 0.0  # no surface in this column: a zero direction misses
        return out


BATCHED = BatchedWorkload()



@dataclass(frozen=True)
class DynamicWorkload:
    """Dynamic world: the small world plus n_edits seeded sphere edits per frame (radius
    uniform in [r_lo, r_hi], half carve / half fill with a material uniform in 1..12), centred
    within 8 voxels of the terrain surface; 1920x1080 primary + sun shadow retrace of the
    dirtied 8x8 blocks; a learned probe update per frame."""
    name: str = "dynamic"
    dims: tuple = (256, 128, 256)
    n_edits: int = 4096
    r_lo: int = 2
    r_hi: int = 12
    width: int = 1920
    height: int = 1080
    seed: int = 0x5EED_0005

    def world(self, threads: int = 0) -> World:
        return World(self.dims).generate(SMALL.seed, threads)

    def camera(self):
        return SMALL.camera()

    def edits(self, frame: int, world: World, n: int | None = None) -> np.ndarray:
        """The edits of one frame (EDIT_DTYPE), applied in array order."""
        n = self.n_edits if n is None else int(n)
        rng = np.random.default_rng([self.seed, int(frame)])
        top = world.top_surface()
        e = np.zeros(n, EDIT_DTYPE)
        xs = rng.integers(0, self.dims[0], n)
        zs = rng.integers(0, self.dims[2], n)
        ys = np.maximum(top[xs, zs], 0) + rng.integers(-8, 9, n)
        e["center"] = np.stack([xs, ys, zs], 1).astype(np.int32)
        e["radius"] = rng.integers(self.r_lo, self.r_hi + 1, n).astype(np.uint8)
        carve = rng.random(n) < 0.5
        e["material"] = np.where(carve, 0, rng.integers(1, 13, n)).astype(np.uint8)
        return e


DYNAMIC = DynamicWorkload()
