"""GI probes on the device: cs_gi_probe_update (shaders/gi_probes.wgsl:274-396).

`ProbeGrid` owns a grid of vg_probe records (GiProbe, gi_probes.wgsl:25-48) on a world's
device and runs the update kernels of libvoxelg_amd.so. The helpers below build the update
parameters (`gi_params`, the reference's constants plus lighting derived from a sun
direction) and evaluate probe irradiance the way the reference's sampler does
(`sh_irradiance`, gi_probes.wgsl:86-95) for error measurement.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (GI_PARAMS_DTYPE, IT_NAMES, PROBE_DTYPE, VG_GI_BOUNCE_TABLE, VG_IT_COUNT,
                   check, f32x3, lib, ptr)

# SH-L1 record fields in the order of the 20 projected coefficients.
SH_FIELDS = ("sh_r", "sh_g", "sh_b", "sh_dist", "sh_dist2")
# Evaluation normals of the irradiance error: the six axis directions.
AXIS_NORMALS = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]],
                        np.float64)


def gi_params(sun, **overrides) -> np.ndarray:
    """vg_gi_params with the reference's defaults and lighting for `sun`, then overrides
    (grid, spacing, rays_per_probe, epochs, update_div, round, bounces, seed, flags, ...)."""
    p = np.zeros(1, GI_PARAMS_DTYPE)
    check(lib().vg_gi_params_default(f32x3(sun), ptr(p)), "vg_gi_params_default")
    for k, v in overrides.items():
        p[k] = v
    return p


def directions(n: int) -> np.ndarray:
    """The n-point spherical-Fibonacci gather set, float32 [n, 4]."""
    out = np.zeros((int(n), 4), np.float32)
    check(lib().vg_gi_directions(int(n), ptr(out)), "vg_gi_directions")
    return out


def bounce_table() -> np.ndarray:
    """The cosine-weighted hemisphere table, float32 [4096, 4] (t, b, cos theta, 0)."""
    out = np.zeros((VG_GI_BOUNCE_TABLE, 4), np.float32)
    check(lib().vg_gi_bounce_table(ptr(out)), "vg_gi_bounce_table")
    return out


def default_palette() -> np.ndarray:
    """default_palette (src/renderer.rs:201-245), float32 [256, 4]."""
    out = np.zeros((256, 4), np.float32)
    check(lib().vg_gi_default_palette(ptr(out)), "vg_gi_default_palette")
    return out


def sh_coefficients(probes: np.ndarray) -> np.ndarray:
    """Displayed SH of vg_probe records as float64 [n, 20] (r, g, b, dist, dist2 x 4)."""
    return np.concatenate([probes[f].astype(np.float64) for f in SH_FIELDS], axis=1)


def sh_irradiance(sh: np.ndarray, normals=AXIS_NORMALS) -> np.ndarray:
    """sh_eval_irradiance (gi_probes.wgsl:86-95): irradiance / pi toward each normal, with
    the reference's deringing cap. sh: [n, >=12] (r, g, b SH-L1). Returns [n, m, 3]."""
    sh = np.asarray(sh, np.float64)
    nrm = np.asarray(normals, np.float64)
    lv = nrm[:, [1, 2, 0]]  # stored coefficient order (c1y, c1z, c1x)
    a1 = 0.488603 * 0.666667
    out = np.zeros((sh.shape[0], nrm.shape[0], 3))
    for c in range(3):
        co = sh[:, 4 * c:4 * c + 4]
        dc = co[:, 0] * 0.282095
        lin = (co[:, 1:4] @ lv.T) * a1
        lmax = np.linalg.norm(co[:, 1:4], axis=1) * a1
        s = np.minimum(1.0, 0.95 * dc / np.maximum(lmax, 1e-6))
        out[:, :, c] = np.maximum(0.0, dc[:, None] + s[:, None] * lin)
    return out


def irradiance_error(sh: np.ndarray, sh_ref: np.ndarray) -> np.ndarray:
    """Per-probe relative irradiance error: sum |E - E_ref| / sum E_ref over the six axis
    normals and three channels (E from sh_irradiance)."""
    e, r = sh_irradiance(sh), sh_irradiance(sh_ref)
    num = np.abs(e - r).reshape(len(e), -1).sum(axis=1)
    den = r.reshape(len(r), -1).sum(axis=1)
    return num / np.maximum(den, 1e-12)


def probe_cells(params, slots, world_origin=(0, 0, 0)) -> np.ndarray:
    """World grid cell of each probe slot (gi_probes.wgsl:281-297), int32 [n, 3]."""
    p = params.reshape(-1)[0]
    gx, gy, gz = (int(v) for v in p["grid"])
    sp = int(p["spacing"])
    slots = np.asarray(slots, np.int64)
    sx, sy, sz = slots % gx, (slots // gx) % gy, slots // (gx * gy)
    ox, oz = int(np.fix(world_origin[0] / sp)), int(np.fix(world_origin[2] / sp))
    g = np.stack([ox + np.mod(sx - ox, gx), sy, oz + np.mod(sz - oz, gz)], axis=1)
    return g.astype(np.int32)


def probe_origins(probes: np.ndarray, params, slots, world_origin=(0, 0, 0)) -> np.ndarray:
    """The ray origin the next update of each slot uses: (g + 0.5) * spacing plus the
    slot's relocation offset (zero for a new slot), in float32 as the device computes it."""
    p = params.reshape(-1)[0]
    g = probe_cells(params, slots, world_origin)
    rec = probes[np.asarray(slots, np.int64)]
    info = rec["info"]
    new = (info[:, 3] < 0.5) | np.any(info[:, :3] != g.astype(np.float32), axis=1)
    off = np.where(new[:, None], np.float32(0), rec["st_info"][:, 1:4]).astype(np.float32)
    base = (g.astype(np.float32) + np.float32(0.5)) * np.float32(int(p["spacing"]))
    return (base + off).astype(np.float32)


class ProbeGrid:
    """vg_probes: the probe records of one grid on a world's device."""

    def __init__(self, world, params: np.ndarray):
        self.world = world
        self.params = params.copy()
        p = params.reshape(-1)[0]
        self.grid = tuple(int(v) for v in p["grid"])
        self.total = self.grid[0] * self.grid[1] * self.grid[2]
        h = C.c_void_p()
        check(lib().vg_probes_create(world._h, ptr(self.params), C.byref(h)), "vg_probes_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().vg_probes_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, params: np.ndarray, stream=None) -> "ProbeGrid":
        p = np.ascontiguousarray(params, GI_PARAMS_DTYPE)
        check(lib().vg_probes_update(self._h, ptr(p), _stream(stream)), "vg_probes_update")
        return self

    def update_counts(self, params: np.ndarray):
        """One update reporting (rays dict: gather/shadow/bounce, walk iteration classes)."""
        p = np.ascontiguousarray(params, GI_PARAMS_DTYPE)
        rays = (C.c_uint64 * 3)()
        cnt = (C.c_uint64 * VG_IT_COUNT)()
        check(lib().vg_probes_update_counts(self._h, ptr(p), rays, cnt), "vg_probes_update_counts")
        return (dict(zip(("gather", "shadow", "bounce"), (int(r) for r in rays))),
                dict(zip(IT_NAMES, (int(c) for c in cnt))))

    def update_learned(self, cache, params: np.ndarray, train: bool = True,
                       stream=None) -> "ProbeGrid":
        """The learned update: gather + sun shadow, the cache's bounce tail, projection and
        fold; with `train`, the cache's online training steps on this update's hits."""
        p = np.ascontiguousarray(params, GI_PARAMS_DTYPE)
        check(lib().vg_probes_update_learned(self._h, ptr(p), cache._h, 1 if train else 0,
                                             _stream(stream)), "vg_probes_update_learned")
        return self

    def update_learned_counts(self, cache, params: np.ndarray, train: bool = True,
                              classes: bool = False):
        """One learned update reporting (rays dict: gather/shadow/bounce, cache queries[,
        walk iteration classes of the path kernel])."""
        p = np.ascontiguousarray(params, GI_PARAMS_DTYPE)
        rays = (C.c_uint64 * 3)()
        nq = C.c_uint64()
        cnt = (C.c_uint64 * VG_IT_COUNT)()
        check(lib().vg_probes_update_learned_counts(self._h, ptr(p), cache._h, 1 if train else 0,
                                                    rays, C.byref(nq), cnt),
              "vg_probes_update_learned_counts")
        out = (dict(zip(("gather", "shadow", "bounce"), (int(r) for r in rays))), int(nq.value))
        return out + (dict(zip(IT_NAMES, (int(c) for c in cnt))),) if classes else out

    def set_timing(self, on: bool = True) -> "ProbeGrid":
        """Record per-launch times of the updates that follow (HIP events on their stream)."""
        check(lib().vg_probes_set_timing(self._h, 1 if on else 0), "vg_probes_set_timing")
        return self

    def timing(self) -> dict:
        """Mean ms per recorded update of each launch (prep, paths, infer, reduce) and the
        number of updates recorded; starts a new record. Synchronous."""
        ms = (C.c_float * 4)()
        n = C.c_int32()
        check(lib().vg_probes_timing(self._h, ms, C.byref(n)), "vg_probes_timing")
        return dict(zip(("prep", "paths", "infer", "reduce"), (float(v) for v in ms)),
                    updates=int(n.value))

    def autotune(self, params: np.ndarray, cache=None, train: bool = False) -> float:
        """Pick the probe path kernels' schedule (waves per CU, refill) on this update; the
        context keeps it. Returns the best time in ms."""
        p = np.ascontiguousarray(params, GI_PARAMS_DTYPE)
        ms = C.c_float()
        check(lib().vg_autotune_probes(self._h, ptr(p), None if cache is None else cache._h,
                                       1 if train else 0, C.byref(ms)), "vg_autotune_probes")
        return float(ms.value)

    def download(self) -> np.ndarray:
        out = np.zeros(self.total, PROBE_DTYPE)
        check(lib().vg_probes_download(self._h, ptr(out)), "vg_probes_download")
        return out

    def upload(self, probes: np.ndarray) -> "ProbeGrid":
        a = np.ascontiguousarray(probes, PROBE_DTYPE)
        assert len(a) == self.total
        check(lib().vg_probes_upload(self._h, ptr(a)), "vg_probes_upload")
        return self

    def reset(self, stream=None) -> "ProbeGrid":
        check(lib().vg_probes_reset(self._h, _stream(stream)), "vg_probes_reset")
        return self

    def frame_indirect(self, params: np.ndarray, camera: np.ndarray, width: int, height: int,
                       hits, out=None, stream=None):
        """indirect_light in probe mode (shaders/rt_gi.wgsl:128-139, sample_probes
        shaders/gi_probes.wgsl:106-179) for every pixel of a traced frame: hits is the
        frame's device hit buffer; returns a device buffer of (r, g, b, 0) float4 per pixel."""
        p = np.ascontiguousarray(params, GI_PARAMS_DTYPE)
        cam = np.ascontiguousarray(camera, np.float32)
        if out is None:
            out = self.world._buf(width * height * 16)
        check(lib().vg_frame_indirect(self._h, ptr(p), ptr(cam), width, height, hits.data_ptr(),
                                      out.data_ptr(), _stream(stream)), "vg_frame_indirect")
        return out

    def set_palette(self, rgba: np.ndarray) -> "ProbeGrid":
        a = np.ascontiguousarray(rgba, np.float32).reshape(256, 4)
        check(lib().vg_probes_set_palette(self._h, ptr(a)), "vg_probes_set_palette")
        return self


def _stream(stream):
    if stream is None:
        return None
    return getattr(stream, "handle", stream)
