"""Device side: a HIP context, device buffers, and a world uploaded to HBM with the
ray-query entry points.

Mirrors the reference's three boundaries:
  raycast(camera_world, dir, world, world_origin) -> Option<PickHit>   src/raycast.rs:21
  the batched Ray[] -> Hit[] query                                     shaders/accel_probe.wgsl:66
  the frame's primary + shadow traces                                  shaders/raymarch.wgsl:576-762
Everything runs through libvoxelg_amd.so (HIP kernels for gfx950); device memory, streams
and events come from the same library, so no other GPU runtime is needed in the process.
Any object with a `data_ptr()` returning a device address on the same device (a torch
tensor, for instance) may be passed where a DeviceBuffer is expected.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (BATCH_DTYPE, EDIT_DTYPE, HIT_DTYPE, IT_NAMES, RAY_DTYPE, VG_IT_COUNT, check,
                   f32x3, i32x3,
                   lib, ptr)
from .world import World


def device_count() -> int:
    n = C.c_int()
    check(lib().vg_device_count(C.byref(n)), "vg_device_count")
    return int(n.value)


@dataclass
class PickHit:
    """src/raycast.rs:13-17: world voxel and the face normal it was entered through."""
    voxel: tuple
    normal: tuple


FACE_NORMALS = {0: (1, 0, 0), 1: (-1, 0, 0), 2: (0, 1, 0), 3: (0, -1, 0), 4: (0, 0, 1),
                5: (0, 0, -1), 6: (0, 0, 0)}


class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().vg_alloc(ctx._h, self.nbytes, C.byref(p)), "vg_alloc")
        self._p = p.value or 0

    def data_ptr(self) -> int:
        return self._p

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().vg_copy_to_device(self.ctx._h, self._p, a.ctypes.data, a.nbytes),
              "vg_copy_to_device")
        return self

    def numpy(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(n, dtype)
        check(lib().vg_copy_to_host(self.ctx._h, out.ctypes.data, self._p, out.nbytes),
              "vg_copy_to_host")
        return out

    def zero(self) -> "DeviceBuffer":
        check(lib().vg_memset(self.ctx._h, self._p, 0, self.nbytes), "vg_memset")
        return self

    def free(self):
        if self._p:
            lib().vg_free(self.ctx._h, self._p)
            self._p = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DevicePtr:
    """A device address inside another allocation (keeps the owner alive)."""

    def __init__(self, owner, offset_bytes: int):
        self.owner = owner
        self._p = owner.data_ptr() + int(offset_bytes)

    def data_ptr(self) -> int:
        return self._p


class Event:
    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        e = C.c_void_p()
        check(lib().vg_event_create(ctx._h, C.byref(e)), "vg_event_create")
        self._h = e.value

    def record(self, stream=None) -> "Event":
        check(lib().vg_event_record(self.ctx._h, self._h, _stream(stream)), "vg_event_record")
        return self

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().vg_event_elapsed_ms(self.ctx._h, self._h, end._h, C.byref(ms)),
              "vg_event_elapsed_ms")
        return float(ms.value)

    def __del__(self):
        try:
            lib().vg_event_destroy(self.ctx._h, self._h)
        except Exception:
            pass


class Stream:
    """A default-priority HIP stream."""

    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        s = C.c_void_p()
        check(lib().vg_stream_create(ctx._h, C.byref(s)), "vg_stream_create")
        self.handle = s.value

    def synchronize(self):
        check(lib().vg_stream_synchronize(self.ctx._h, self.handle), "vg_stream_synchronize")

    def __del__(self):
        try:
            lib().vg_stream_destroy(self.ctx._h, self.handle)
        except Exception:
            pass


def _stream(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, Stream):
        return stream.handle
    h = getattr(stream, "cuda_stream", None)  # a torch stream
    return h if h is not None else int(stream)


def mask_words(width: int, height: int) -> int:
    """u32 words of a frame's 8x8-block mask (vg_frame_mask_words)."""
    n = C.c_uint32()
    check(lib().vg_frame_mask_words(width, height, C.byref(n)), "vg_frame_mask_words")
    return int(n.value)


def decode_hits(hits, count: int | None = None) -> np.ndarray:
    """Device hits (DeviceBuffer, torch tensor or numpy) -> numpy structured HIT_DTYPE."""
    if isinstance(hits, DeviceBuffer):
        return hits.numpy(HIT_DTYPE, count)
    if hasattr(hits, "detach"):
        hits = hits.detach().cpu().numpy()
    return np.ascontiguousarray(hits).view(HIT_DTYPE).reshape(-1)


class Context:
    """One HIP device (vg_ctx)."""

    def __init__(self, device: int = 0):
        self.device = device
        h = C.c_void_p()
        check(lib().vg_ctx_create(device, C.byref(h)), "vg_ctx_create")
        self._h = h

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, a: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(a)
        return DeviceBuffer(self, max(a.nbytes, 1)).upload(a)

    def rays(self, rays: np.ndarray) -> DeviceBuffer:
        return self.to_device(np.ascontiguousarray(rays, RAY_DTYPE))

    def synchronize(self):
        check(lib().vg_synchronize(self._h), "vg_synchronize")

    def stream(self) -> Stream:
        return Stream(self)

    def event(self) -> Event:
        return Event(self)

    def world(self, dims) -> "DeviceWorld":
        return DeviceWorld(self, dims)

    def set_option(self, option: int, value: int) -> "Context":
        check(lib().vg_ctx_set_option(self._h, option, int(value)), "vg_ctx_set_option")
        return self

    def get_option(self, option: int) -> int:
        v = C.c_int()
        check(lib().vg_ctx_get_option(self._h, option, C.byref(v)), "vg_ctx_get_option")
        return int(v.value)

    def close(self):
        if self._h:
            lib().vg_ctx_destroy(self._h)
            self._h = None


class DeviceWorld:
    """A world's HBM copy (vg_world)."""

    def __init__(self, ctx: Context, dims):
        self.ctx = ctx
        self.dims = tuple(int(d) for d in dims)
        self.origin = (0, 0, 0)
        h = C.c_void_p()
        check(lib().vg_world_create(ctx._h, i32x3(self.dims), C.byref(h)), "vg_world_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().vg_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, world: World) -> "DeviceWorld":
        if tuple(world.dims) != self.dims:
            raise ValueError(f"world dims {world.dims} != device world dims {self.dims}")
        check(lib().vg_world_upload(self._h, ptr(world.bricks), ptr(world.tile_mask),
                                    ptr(world.chunk_mask), ptr(world.l4_mask),
                                    ptr(world.brick_uniform), ptr(world.tile_uniform),
                                    i32x3(world.origin)), "vg_world_upload")
        self.origin = tuple(world.origin)
        return self

    def generate(self, seed: int, origin=(0, 0, 0), stream=None) -> "DeviceWorld":
        """The benchmark terrain generated on the device (the bytes World.generate +
        upload would give), pyramid and uniform tables included."""
        check(lib().vg_world_generate(self._h, int(seed) & 0xFFFFFFFFFFFFFFFF, i32x3(origin),
                                      _stream(stream)), "vg_world_generate")
        self.origin = tuple(origin)
        return self

    # ------------------------------------------------------------ queries --
    def _buf(self, nbytes):
        return self.ctx.alloc(max(int(nbytes), 1))

    def trace_rays(self, rays, n: int | None = None, flags: int = 0, hits=None, stream=None):
        """Batched query. rays: DeviceBuffer of vg_ray (or a numpy RAY_DTYPE array, copied)."""
        if isinstance(rays, np.ndarray):
            n = len(rays) if n is None else n
            rays = self.ctx.rays(rays)
        n = rays.nbytes // 32 if n is None else int(n)
        if hits is None:
            hits = self._buf(n * 20)
        check(lib().vg_trace_rays(self._h, rays.data_ptr(), n, flags, hits.data_ptr(),
                                  _stream(stream)), "vg_trace_rays")
        return hits

    def trace_rays_counts(self, rays, n: int | None = None, flags: int = 0):
        if isinstance(rays, np.ndarray):
            n = len(rays) if n is None else n
            rays = self.ctx.rays(rays)
        n = rays.nbytes // 32 if n is None else int(n)
        hits = self._buf(n * 20)
        counts = (C.c_uint64 * VG_IT_COUNT)()
        check(lib().vg_trace_rays_counts(self._h, rays.data_ptr(), n, flags, hits.data_ptr(),
                                         counts), "vg_trace_rays_counts")
        return hits, dict(zip(IT_NAMES, (int(c) for c in counts)))

    def trace_primary(self, camera: np.ndarray, width: int, height: int, tmax: float,
                      flags: int = 0, hits=None, stream=None):
        cam = np.ascontiguousarray(camera, np.float32)
        if hits is None:
            hits = self._buf(width * height * 20)
        check(lib().vg_trace_primary(self._h, ptr(cam), width, height, tmax, flags,
                                     hits.data_ptr(), _stream(stream)), "vg_trace_primary")
        return hits

    def trace_shadow(self, camera: np.ndarray, width: int, height: int, primary_hits, sun,
                     tmax: float, occluded=None, stream=None):
        cam = np.ascontiguousarray(camera, np.float32)
        if occluded is None:
            occluded = self._buf(width * height)
        check(lib().vg_trace_shadow(self._h, ptr(cam), width, height, primary_hits.data_ptr(),
                                    f32x3(sun), tmax, occluded.data_ptr(), _stream(stream)),
              "vg_trace_shadow")
        return occluded

    def trace_frame(self, camera: np.ndarray, width: int, height: int, primary_tmax: float,
                    primary_flags: int, sun, shadow_tmax: float, hits=None, occluded=None,
                    stream=None):
        cam = np.ascontiguousarray(camera, np.float32)
        if hits is None:
            hits = self._buf(width * height * 20)
        if occluded is None:
            occluded = self._buf(width * height)
        check(lib().vg_trace_frame(self._h, ptr(cam), width, height, primary_tmax, primary_flags,
                                   f32x3(sun), shadow_tmax, hits.data_ptr(), occluded.data_ptr(),
                                   _stream(stream)), "vg_trace_frame")
        return hits, occluded

    def trace_water_secondary(self, camera: np.ndarray, width: int, height: int, hits,
                              refl=None, refr=None, stream=None):
        """The reflection and refraction rays of every water pixel of a traced frame
        (shade_water_top, shaders/raymarch.wgsl:3534-3667), traced with
        VG_RAY_SKIP_TRANSPARENT and tmax 200: (refl, refr) device hit buffers, one record
        per pixel (misses where the primary hit is not water)."""
        cam = np.ascontiguousarray(camera, np.float32)
        if refl is None:
            refl = self._buf(width * height * 20)
        if refr is None:
            refr = self._buf(width * height * 20)
        check(lib().vg_trace_water_secondary(self._h, ptr(cam), width, height, hits.data_ptr(),
                                             refl.data_ptr(), refr.data_ptr(), _stream(stream)),
              "vg_trace_water_secondary")
        return refl, refr

    # ------------------------------------------------------------ dynamic --
    def apply_edits(self, edits: np.ndarray, stream=None) -> "DeviceWorld":
        """Sphere edits (EDIT_DTYPE records, applied in order) with the incremental pyramid
        and uniform-table rebuild on the device."""
        e = np.ascontiguousarray(edits, EDIT_DTYPE)
        check(lib().vg_world_apply_edits(self._h, ptr(e) if len(e) else None, len(e),
                                         _stream(stream)), "vg_world_apply_edits")
        return self

    def dirty_blocks(self, camera: np.ndarray, width: int, height: int, sun=None,
                     shadow_tmax: float = 480.0, mask=None, stream=None):
        """Device block mask of the 8x8 blocks the last edits can change (sun None: primary
        rays only); into `mask` (mask_words(width, height) u32) when given."""
        cam = np.ascontiguousarray(camera, np.float32)
        if mask is None:
            mask = self._buf(mask_words(width, height) * 4)
        check(lib().vg_frame_dirty_blocks(self._h, ptr(cam), width, height,
                                          None if sun is None else f32x3(sun), shadow_tmax,
                                          mask.data_ptr(), _stream(stream)),
              "vg_frame_dirty_blocks")
        return mask

    def trace_frame_blocks(self, camera: np.ndarray, width: int, height: int,
                           primary_tmax: float, primary_flags: int, sun, shadow_tmax: float,
                           mask, hits, occluded=None, stream=None):
        """Retrace the masked 8x8 blocks into existing hit / shadow buffers."""
        cam = np.ascontiguousarray(camera, np.float32)
        check(lib().vg_trace_frame_blocks(self._h, ptr(cam), width, height, primary_tmax,
                                          primary_flags, None if sun is None else f32x3(sun),
                                          shadow_tmax, mask.data_ptr(), hits.data_ptr(),
                                          None if occluded is None else occluded.data_ptr(),
                                          _stream(stream)), "vg_trace_frame_blocks")
        return hits, occluded

    def download(self) -> World:
        """The device state back in the reference layout, as a World."""
        out = World(self.dims, self.origin)
        check(lib().vg_world_download(self._h, ptr(out.bricks), ptr(out.tile_mask),
                                      ptr(out.chunk_mask), ptr(out.l4_mask),
                                      ptr(out.brick_uniform), ptr(out.tile_uniform)),
              "vg_world_download")
        return out

    def trace_frame_counts(self, camera: np.ndarray, width: int, height: int,
                           primary_tmax: float, primary_flags: int, sun, shadow_tmax: float):
        cam = np.ascontiguousarray(camera, np.float32)
        hits = self._buf(width * height * 20)
        occluded = self._buf(width * height)
        counts = (C.c_uint64 * VG_IT_COUNT)()
        check(lib().vg_trace_frame_counts(self._h, ptr(cam), width, height, primary_tmax,
                                          primary_flags, f32x3(sun), shadow_tmax,
                                          hits.data_ptr(), occluded.data_ptr(), counts),
              "vg_trace_frame_counts")
        return hits, occluded, dict(zip(IT_NAMES, (int(c) for c in counts)))

    def autotune_frame(self, camera: np.ndarray, width: int, height: int, primary_tmax: float,
                       primary_flags: int, sun, shadow_tmax: float) -> float:
        """Time every (waves per CU, refill) pair of the persistent frame kernel on this frame
        and keep the fastest in the context; returns its time in ms."""
        cam = np.ascontiguousarray(camera, np.float32)
        hits = self._buf(width * height * 20)
        occ = self._buf(width * height)
        ms = C.c_float()
        check(lib().vg_autotune_frame(self._h, ptr(cam), width, height, primary_tmax,
                                      primary_flags, f32x3(sun), shadow_tmax, hits.data_ptr(),
                                      occ.data_ptr(), C.byref(ms)), "vg_autotune_frame")
        return float(ms.value)

    def autotune_rays(self, rays, n: int | None = None, flags: int = 0) -> float:
        if isinstance(rays, np.ndarray):
            n = len(rays) if n is None else n
            rays = self.ctx.rays(rays)
        n = rays.nbytes // 32 if n is None else int(n)
        hits = self._buf(n * 20)
        ms = C.c_float()
        check(lib().vg_autotune_rays(self._h, rays.data_ptr(), n, flags, hits.data_ptr(),
                                     C.byref(ms)), "vg_autotune_rays")
        return float(ms.value)

    # ------------------------------------------------------------ batches --
    def batch_set(self, batches) -> "BatchSet":
        """Device descriptors for many uneven batches: a list of (rays, hits, n, flags) with
        rays / hits DeviceBuffers (or anything with data_ptr())."""
        return BatchSet(self, batches)

    def trace_batches(self, bs: "BatchSet", stream=None):
        check(lib().vg_trace_batches(self._h, bs.desc.data_ptr(), bs.prefix.data_ptr(), bs.n,
                                     bs.total, _stream(stream)), "vg_trace_batches")

    def trace_batches_counts(self, bs: "BatchSet") -> dict:
        counts = (C.c_uint64 * VG_IT_COUNT)()
        check(lib().vg_trace_batches_counts(self._h, bs.desc.data_ptr(), bs.prefix.data_ptr(),
                                            bs.n, bs.total, counts), "vg_trace_batches_counts")
        return dict(zip(IT_NAMES, (int(c) for c in counts)))

    def autotune_batches(self, bs: "BatchSet") -> float:
        ms = C.c_float()
        check(lib().vg_autotune_batches(self._h, bs.desc.data_ptr(), bs.prefix.data_ptr(), bs.n,
                                        bs.total, C.byref(ms)), "vg_autotune_batches")
        return float(ms.value)

    def raycast_hit(self, camera_world, direction) -> np.void:
        out = np.zeros(1, HIT_DTYPE)
        check(lib().vg_raycast(self._h, f32x3(camera_world), f32x3(direction), ptr(out)),
              "vg_raycast")
        return out[0]


def raycast(camera_world, direction, world: DeviceWorld, world_origin=None):
    """src/raycast.rs:21 `raycast(camera_world, dir, &World, world_origin) -> Option<PickHit>`,
    traced on the GPU against the uploaded window. world_origin must match the upload."""
    if world_origin is not None and tuple(int(v) for v in world_origin) != world.origin:
        raise ValueError(f"world_origin {tuple(world_origin)} != uploaded origin {world.origin}")
    h = world.raycast_hit(camera_world, direction)
    if not h["hit"]:
        return None
    return PickHit(tuple(int(v) for v in h["voxel"]), FACE_NORMALS[int(h["face"])])


class BatchSet:
    """vg_batch descriptors and their prefix sums, resident on the device."""

    def __init__(self, world: DeviceWorld, batches):
        d = np.zeros(len(batches), BATCH_DTYPE)
        sizes = []
        for i, (rays, hits, n, flags) in enumerate(batches):
            d[i] = (rays.data_ptr(), hits.data_ptr(), int(n), int(flags))
            sizes.append(int(n))
        prefix = np.zeros(len(batches) + 1, np.uint32)
        prefix[1:] = np.cumsum(sizes, dtype=np.uint64).astype(np.uint32)
        self.n = len(batches)
        self.total = int(prefix[-1])
        self.sizes = sizes
        self._keep = batches
        self.desc = world.ctx.to_device(d)
        self.prefix = world.ctx.to_device(prefix)


class Graph:
    """Launches captured from a stream into an executable HIP graph."""

    def __init__(self, ctx: Context, stream: Stream, fn):
        check(lib().vg_graph_begin(ctx._h, stream.handle), "vg_graph_begin")
        try:
            fn(stream)
        finally:
            g = C.c_void_p()
            st = lib().vg_graph_end(ctx._h, stream.handle, C.byref(g))
        check(st, "vg_graph_end")
        self._g = g
        self.ctx = ctx

    def launch(self, stream: Stream):
        check(lib().vg_graph_launch(self._g, stream.handle), "vg_graph_launch")

    def __del__(self):
        try:
            lib().vg_graph_destroy(self._g)
        except Exception:
            pass

