"""CPU ORACLE — test infrastructure only.

ctypes wrapper over oracle/liboracle.so (oracle.c: a plain-C restatement of voxelG's ray
query, see oracle.h for what each function follows and what pins it). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and only as
the checker or the timed CPU baseline; the product (voxelg_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

BRICK_DTYPE = np.dtype([("occupancy", "<u8"), ("materials", "u1", (64,))])
RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("tmax", "<f4"), ("d", "<f4", (3,)), ("flags", "<u4")])
HIT_DTYPE = np.dtype(
    [("voxel", "<i4", (3,)), ("t", "<f4"), ("face", "u1"), ("material", "u1"), ("hit", "u1"),
     ("pad", "u1")]
)
LOD, SKIP_TRANSPARENT, ANY_HIT, BEAM = 1, 2, 4, 8
NO_GROUPS = 0x100  # oracle only: the reference's levels without 2x2x2 group skips
NO_YTOP = 0x200  # oracle only: no y-bound exit (orc_world.ytop)
FACE_NONE = 6
IT_COUNT = 14
IT_NAMES = ("l4_skip", "chunk_skip", "tile_skip", "tile_lod", "tile_uniform", "brick_skip",
            "brick_uniform", "brick_lod", "voxel", "chunk_group_skip", "tile_group_skip",
            "brick_group_skip", "voxel_group_skip", "beam_test")
# Algorithmic bytes per iteration class: the pyramid words the walk's test chain reads in
# the reference's order (L4 word 8, chunk word 8, tile-uniform byte 1, tile word 8,
# brick-uniform byte 1, occupancy word 8); a tile-LOD test adds its tile word, the first
# child brick's occupancy and one material byte (17) to the class that follows it; a brick
# LOD reads the occupancy and one material byte past the brick-uniform byte.
IT_BYTES = (8, 8, 16, 17, 17, 25, 26, 35, 34, 8, 16, 25, 34, 8)


# orc_gi_params (oracle.h): the byte layout of vg_gi_params.
GI_PARAMS_DTYPE = np.dtype([
    ("grid", "<i4", (3,)), ("spacing", "<i4"), ("rays_per_probe", "<i4"), ("epochs", "<i4"),
    ("update_div", "<i4"), ("round", "<i4"), ("bounces", "<i4"), ("seed", "<u4"),
    ("flags", "<u4"), ("hysteresis", "<f4"), ("gi_dist", "<f4"), ("shadow_tmax", "<f4"),
    ("sun", "<f4", (3,)), ("gi_sky", "<f4"), ("sun_color", "<f4", (3,)),
    ("ambient_tail", "<f4"), ("ambient", "<f4", (3,)), ("day_t", "<f4"),
    ("sky_horizon", "<f4", (3,)), ("dusk_w", "<f4"), ("sky_zenith", "<f4", (3,)),
    ("halo", "<f4"), ("sun_xz", "<f4", (2,)), ("pad", "<f4", (2,))])
# orc_nrc_config / orc_nrc_query (oracle.h): the byte layouts of vg_nrc_config / vg_nrc_query.
NRC_CONFIG_DTYPE = np.dtype([
    ("levels", "<i4"), ("log2_table", "<i4"), ("scale", "<f4", (16,)), ("lr_grid", "<f4"),
    ("lr_mlp", "<f4"), ("beta1", "<f4"), ("beta2", "<f4"), ("eps", "<f4"), ("init_grid", "<f4"),
    ("seed", "<u4"), ("batch", "<i4"), ("steps", "<i4"), ("pad", "<i4")])
NRC_QUERY_DTYPE = np.dtype([("pos", "<f4", (3,)), ("face", "<u4"), ("dir", "<f4", (3,)),
                            ("aux", "<u4")])
NRC_WIDTH = 64
NRC_MLP_PARAMS = 3 * 64 * 64 + 16 * 64
NRC_IN_FACE, NRC_IN_DIR, NRC_IN_ONE = 32, 38, 41
EDIT_DTYPE = np.dtype([("center", "<i4", (3,)), ("radius", "u1"), ("material", "u1"),
                       ("pad", "u1", (2,))])
GI_SKIP_BURIED = 1
GI_BOUNCE_TABLE = 4096
PROBE_FIELDS = 48


class OrcWorld(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("origin", C.c_int32 * 3), ("bricks", C.c_void_p),
                ("tile_mask", C.c_void_p), ("chunk_mask", C.c_void_p), ("l4_mask", C.c_void_p),
                ("brick_uniform", C.c_void_p), ("tile_uniform", C.c_void_p), ("ytop", C.c_int32)]


def world_ytop(dims, origin, tile_mask, tile_uniform) -> int:
    """orc_world.ytop of a world's arrays: 4 x (the highest window-local brick row holding a
    set tile bit + 1), a uniform tile counting as all its rows; dims[1] when the window folds
    y. Tiles are in the reference order (x fastest, then y, then z)."""
    if int(origin[1]) % int(dims[1]) != 0:
        return int(dims[1])
    t0, t1 = int(dims[0]) // 16, int(dims[1]) // 16
    m = np.asarray(tile_mask, np.uint64)
    ty = (np.arange(len(m), dtype=np.int64) // t0) % t1
    by = np.where(m >> np.uint64(48), 3, np.where(m >> np.uint64(32), 2,
                                                   np.where(m >> np.uint64(16), 1, 0)))
    top = np.where(np.asarray(tile_uniform) != 0, ty * 16 + 16,
                   np.where(m != 0, ty * 16 + by * 4 + 4, 0))
    return int(top.max()) if len(top) else 0


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(path):
    """The oracle library at path with every entry point's signature set."""
    L = C.CDLL(path)
    P, I32P, U64P, F32P = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_uint64), C.POINTER(C.c_float)
    W = C.POINTER(OrcWorld)
    sig = {
        "orc_counts": (None, [I32P, C.POINTER(C.c_int64)]),
        "orc_brick_uniform_of": (C.c_uint8, [P]),
        "orc_rebuild_all_masks": (None, [W]),
        "orc_rebuild_all_uniform": (None, [W]),
        "orc_set_voxel": (None, [W, C.c_int32, C.c_int32, C.c_int32, C.c_uint8]),
        "orc_apply_spheres": (None, [W, P, C.c_int64]),
        "orc_material_at_world": (C.c_uint8, [W, C.c_int32, C.c_int32, C.c_int32]),
        "orc_hash2": (C.c_float, [C.c_int32, C.c_int32]),
        "orc_hash3": (C.c_float, [C.c_int32, C.c_int32, C.c_int32]),
        "orc_value_noise_2d": (C.c_float, [C.c_float, C.c_float]),
        "orc_value_noise_3d": (C.c_float, [C.c_float, C.c_float, C.c_float]),
        "orc_fbm_2d": (C.c_float, [C.c_float, C.c_float, C.c_uint32]),
        "orc_seed_offset_xz": (None, [C.c_uint64, F32P]),
        "orc_terrain_height": (C.c_int32, [I32P, C.c_uint64, C.c_int32, C.c_int32]),
        "orc_generate": (None, [W, C.c_uint64]),
        "orc_raycast_r0": (None, [W, F32P, F32P, C.c_float, P]),
        "orc_trace_r1": (None, [W, P, C.c_uint32, P, U64P]),
        "orc_trace_flat": (None, [W, P, C.c_uint32, P]),
        "orc_trace_brute": (None, [W, P, P, I32P]),
        "orc_trace_r1_batch": (None, [W, P, C.c_size_t, C.c_uint32, P, U64P, C.c_int]),
        "orc_raycast_r0_batch": (None, [W, P, C.c_size_t, P, C.c_int]),
        "orc_primary_dir": (None, [P, C.c_int32, C.c_int32, F32P]),
        "orc_primary_rays": (None, [P, C.c_int32, C.c_int32, C.c_float, C.c_uint32, P]),
        "orc_frame": (None, [W, P, C.c_int32, C.c_int32, C.c_float, C.c_uint32, F32P,
                             C.c_float, C.c_int32, C.c_int32, C.c_int32, P, P, C.c_int, U64P,
                             P]),
        "orc_frame_r0": (None, [W, P, C.c_int32, C.c_int32, C.c_float, F32P, C.c_float,
                                C.c_int32, C.c_int32, C.c_int32, P, P, C.c_int]),
        "orc_shadow_origin": (None, [F32P, F32P, P, F32P]),
        "orc_trace_dir": (C.c_int, [W, F32P, F32P, C.c_float, C.c_uint32, P, U64P]),
        "orc_gi_sky": (None, [P, F32P, C.POINTER(C.c_double)]),
        "orc_gi_bounce_index": (C.c_uint32, [C.c_uint32, C.c_int32, C.c_uint32, C.c_uint32]),
        "orc_gi_sample": (None, [W, P, P, P, F32P, F32P, C.c_uint32, C.c_uint32,
                                 C.POINTER(C.c_double), C.POINTER(C.c_uint8), F32P, U64P]),
        "orc_gi_update": (None, [W, P, P, P, P, P, C.c_int, U64P]),
        "orc_gi_reference": (None, [W, P, C.c_int32, P, P, P, P, P, C.c_int64, P, C.c_int]),
        "orc_gi_update_learned": (None, [W, P, P, P, P, P, P, C.c_int, U64P]),
        "orc_gi_bounce_tail": (None, [W, P, P, P, P, C.c_uint32, P, U64P]),
        "orc_gi_bounce_tails": (None, [W, P, P, P, C.c_int64, P, P, P, C.c_int]),
        "orc_water_rays": (None, [P, C.c_int32, C.c_int32, P, P]),
        "orc_sh_irradiance": (None, [P, F32P, F32P]),
        "orc_sample_probes": (None, [P, I32P, C.c_int32, F32P, F32P, F32P, F32P]),
        "orc_frame_indirect": (None, [P, I32P, C.c_int32, P, C.c_int32, C.c_int32, P, P]),
        "orc_bf16_round": (C.c_double, [C.c_double]),
        "orc_nrc_hash": (C.c_uint32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
        "orc_nrc_scales": (None, [C.c_int32, C.c_float, C.c_float, P]),
        "orc_nrc_encode": (None, [P, P, P, C.c_int, P]),
        "orc_nrc_forward": (None, [P, P, P, C.c_int64, P, C.c_int, P]),
        "orc_nrc_loss": (C.c_double, [P, P, P, C.c_int64, P, P, C.c_int]),
        "orc_nrc_loss_grad": (C.c_double, [P, P, P, C.c_int64, P, P, C.c_int, P, P]),
        "orc_nrc_adam": (None, [C.c_int64, P, P, P, P, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int64, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


_timing = None
TIMING_FLAGS = ["-O3", "-march=native", "-std=c11", "-fPIC", "-ffp-contract=off",
                "-fno-fast-math", "-pthread", "-D_GNU_SOURCE"]


def timing_lib():
    """(library, description): the same sources built for timing on the host that runs
    this (gcc -O3 -march=native, contraction off: the results are those of the test build),
    into a private temporary directory; the test build (-O2) when no compiler is present."""
    global _timing
    if _timing is None:
        import shutil
        import tempfile
        cc = shutil.which("gcc") or shutil.which("cc")
        if cc:
            out = os.path.join(tempfile.mkdtemp(prefix="oracle_timing_"), "liboracle_timing.so")
            src = [os.path.join(_HERE, f) for f in ("oracle.c", "oracle_gi.c", "oracle_nrc.c")]
            r = subprocess.run([cc] + TIMING_FLAGS + ["-shared", "-o", out] + src + ["-lm"],
                               capture_output=True, text=True)
            if r.returncode == 0:
                _timing = (_load(out), os.path.basename(cc) + " " + " ".join(TIMING_FLAGS[:2]))
        if _timing is None:
            _timing = (lib(), "test build (gcc -O2)")
    return _timing


def host_cpu() -> dict:
    """The host's CPU model, its physical cores and logical CPUs (Linux /proc/cpuinfo),
    and the CPUs this process may run on."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k:
                if core is not None:
                    cores.add((phys, core))
                phys = core = None
        if core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "physical_cores": len(cores) or None,
            "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def _threads(threads: int) -> int:
    """Worker threads: the request, else the host's cores capped at 16 (one GPU's CPU share)."""
    return int(threads) if threads else min(16, os.cpu_count() or 1)


def _rows(rows, height):
    if rows is None:
        return 0, int(height), 1
    r0, r1 = int(rows[0]), int(rows[1])
    return r0, r1, int(rows[2]) if len(rows) > 2 else 1


def _f3(v):
    return (C.c_float * 3)(*[float(np.float32(x)) for x in v])


def _i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def counts(dims):
    out = (C.c_int64 * 5)()
    lib().orc_counts(_i3(dims), out)
    return tuple(int(x) for x in out)


class OracleWorld:
    """A world in the reference layout whose arrays the oracle reads (and may write).

    `OracleWorld.view(w)` wraps any object with the World attributes (dims, origin, bricks,
    tile_mask, chunk_mask, l4_mask, brick_uniform, tile_uniform) without copying."""

    def __init__(self, dims, origin=(0, 0, 0)):
        nb, nt, nc, nl4, _ = counts(dims)
        self.dims = tuple(int(d) for d in dims)
        self.origin = tuple(int(o) for o in origin)
        self.bricks = np.zeros(nb, BRICK_DTYPE)
        self.tile_mask = np.zeros(nt, np.uint64)
        self.chunk_mask = np.zeros(nc, np.uint64)
        self.l4_mask = np.zeros(max(nl4, 1), np.uint64)
        self.brick_uniform = np.zeros(nb, np.uint8)
        self.tile_uniform = np.zeros(nt, np.uint8)

    @classmethod
    def view(cls, w) -> "OracleWorld":
        o = cls.__new__(cls)
        for a in ("dims", "origin", "bricks", "tile_mask", "chunk_mask", "l4_mask",
                  "brick_uniform", "tile_uniform"):
            setattr(o, a, getattr(w, a))
        return o

    @classmethod
    def copy_of(cls, w) -> "OracleWorld":
        """An independent copy of a world's arrays."""
        o = cls.__new__(cls)
        o.dims, o.origin = tuple(w.dims), tuple(w.origin)
        for a in ("bricks", "tile_mask", "chunk_mask", "l4_mask", "brick_uniform",
                  "tile_uniform"):
            setattr(o, a, np.array(getattr(w, a), copy=True))
        return o

    def cstruct(self) -> OrcWorld:
        s = OrcWorld()
        s.dims = _i3(self.dims)
        s.origin = _i3(self.origin)
        s.bricks = _p(self.bricks)
        s.tile_mask = _p(self.tile_mask)
        s.chunk_mask = _p(self.chunk_mask)
        s.l4_mask = _p(self.l4_mask)
        s.brick_uniform = _p(self.brick_uniform)
        s.tile_uniform = _p(self.tile_uniform)
        s.ytop = world_ytop(self.dims, self.origin, self.tile_mask, self.tile_uniform)
        self._c = s  # keep alive
        return s

    # -------------------------------------------------------------- build --
    def generate(self, seed: int) -> "OracleWorld":
        lib().orc_generate(C.byref(self.cstruct()), C.c_uint64(seed))
        return self

    def rebuild(self) -> "OracleWorld":
        s = C.byref(self.cstruct())
        lib().orc_rebuild_all_masks(s)
        lib().orc_rebuild_all_uniform(s)
        return self

    def set_voxel(self, x, y, z, mat) -> None:
        """Reference set_voxel on WINDOW-LOCAL slot coordinates (incremental masks)."""
        lib().orc_set_voxel(C.byref(self.cstruct()), int(x), int(y), int(z), int(mat))

    def apply_edit(self, wx, wy, wz, mat) -> None:
        """World coordinates -> toroidal slot, then set_voxel (src/voxel.rs:450-474)."""
        rel = [wx - self.origin[0], wy - self.origin[1], wz - self.origin[2]]
        if any(r < 0 or r >= d for r, d in zip(rel, self.dims)):
            return
        s = [int(np.mod(v, d)) for v, d in zip((wx, wy, wz), self.dims)]
        self.set_voxel(s[0], s[1], s[2], mat)

    def apply_spheres(self, edits: np.ndarray) -> "OracleWorld":
        """apply_sphere for each edit in order, then a full pyramid / LUT rebuild."""
        e = np.ascontiguousarray(edits, EDIT_DTYPE)
        lib().orc_apply_spheres(C.byref(self.cstruct()), _p(e), len(e))
        return self

    def material_at(self, wx, wy, wz) -> int:
        return int(lib().orc_material_at_world(C.byref(self.cstruct()), int(wx), int(wy), int(wz)))

    # ------------------------------------------------------------- traces --
    def raycast_r0(self, cam, d, tmax=np.inf) -> np.void:
        out = np.zeros(1, HIT_DTYPE)
        lib().orc_raycast_r0(C.byref(self.cstruct()), _f3(cam), _f3(d), float(tmax), _p(out))
        return out[0]

    def trace_r1(self, o, d, tmax=np.inf, flags=0):
        ray = make_rays([o], [d], tmax, 0)
        out = np.zeros(1, HIT_DTYPE)
        cnt = (C.c_uint64 * IT_COUNT)()
        lib().orc_trace_r1(C.byref(self.cstruct()), _p(ray), flags, _p(out), cnt)
        return out[0], dict(zip(IT_NAMES, (int(c) for c in cnt)))

    def trace_flat(self, rays: np.ndarray, flags=0) -> np.ndarray:
        out = np.zeros(len(rays), HIT_DTYPE)
        s = C.byref(self.cstruct())
        for i in range(len(rays)):
            lib().orc_trace_flat(s, rays[i:i + 1].ctypes.data, flags, out[i:i + 1].ctypes.data)
        return out

    def trace_brute(self, rays: np.ndarray):
        out = np.zeros(len(rays), HIT_DTYPE)
        ties = np.zeros(len(rays), np.int32)
        s = C.byref(self.cstruct())
        t = C.c_int32()
        for i in range(len(rays)):
            lib().orc_trace_brute(s, rays[i:i + 1].ctypes.data, out[i:i + 1].ctypes.data,
                                  C.byref(t))
            ties[i] = t.value
        return out, ties

    def trace_rays(self, rays: np.ndarray, flags=0, threads=0, L=None):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(len(rays), HIT_DTYPE)
        cnt = (C.c_uint64 * IT_COUNT)()
        (L or lib()).orc_trace_r1_batch(C.byref(self.cstruct()), _p(rays), len(rays), flags,
                                        _p(out),
                                 cnt, _threads(threads))
        return out, dict(zip(IT_NAMES, (int(c) for c in cnt)))

    def water_secondary(self, camera, width, height, hits: np.ndarray, threads=0):
        """Reflection and refraction hit records per pixel of a frame's primary hits: the
        rays of orc_water_rays traced by R1 (misses where the primary hit is not water)."""
        rays = water_rays(camera, width, height, hits)
        out, _ = self.trace_rays(rays, threads=threads)
        out = out.reshape(height * width, 2)
        return np.ascontiguousarray(out[:, 0]), np.ascontiguousarray(out[:, 1])

    def raycast_rays(self, rays: np.ndarray, threads=0, L=None) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(len(rays), HIT_DTYPE)
        (L or lib()).orc_raycast_r0_batch(C.byref(self.cstruct()), _p(rays), len(rays), _p(out),
                                   _threads(threads))
        return out

    def frame(self, camera, width, height, primary_tmax, primary_flags, sun, shadow_tmax,
              rows=None, threads=0, iters=False, L=None):
        """R1 frame: hits (+ shadow bytes) for rows (r0, r1) or (r0, r1, step) of the image,
        i.e. r0, r0 + step, ... below r1; with iters=True also the per-pixel (primary,
        shadow) iteration counts. L: another build of the oracle library."""
        r0, r1, st = _rows(rows, height)
        n = len(range(r0, r1, st)) * width
        hits = np.zeros(n, HIT_DTYPE)
        shadow = np.zeros(n, np.uint8)
        it = np.zeros((n, 2), np.uint32) if iters else None
        cnt = (C.c_uint64 * IT_COUNT)()
        cam = np.ascontiguousarray(camera, np.float32)
        (L or lib()).orc_frame(C.byref(self.cstruct()), _p(cam), width, height, primary_tmax,
                               primary_flags, _f3(sun if sun is not None else (0, 0, 0)),
                               shadow_tmax, r0, r1, st, _p(hits), _p(shadow), _threads(threads),
                               cnt, _p(it) if iters else None)
        out = (hits, shadow, dict(zip(IT_NAMES, (int(c) for c in cnt))))
        return out + (it,) if iters else out

    def frame_r0(self, camera, width, height, primary_tmax, sun, shadow_tmax, rows=None,
                 threads=0, L=None):
        """The same frame traced by R0 (raycast() of src/raycast.rs per pixel, then per hit)."""
        r0, r1, st = _rows(rows, height)
        n = len(range(r0, r1, st)) * width
        hits = np.zeros(n, HIT_DTYPE)
        shadow = np.zeros(n, np.uint8)
        cam = np.ascontiguousarray(camera, np.float32)
        (L or lib()).orc_frame_r0(C.byref(self.cstruct()), _p(cam), width, height, primary_tmax,
                                  _f3(sun if sun is not None else (0, 0, 0)), shadow_tmax, r0,
                                  r1, st, _p(hits), _p(shadow), _threads(threads))
        return hits, shadow

    # ---------------------------------------------------------- GI probes --
    def trace_dir(self, o, d, tmax, flags):
        """R1 along d exactly as given; returns (started, hit)."""
        out = np.zeros(1, HIT_DTYPE)
        started = lib().orc_trace_dir(C.byref(self.cstruct()), _f3(o), _f3(d), float(tmax),
                                      int(flags), _p(out), None)
        return bool(started), out[0]

    def gi_sample(self, params, btab, palette, o, d, probe, k):
        """One gather sample: (rgb + visible distance as float64[4], face, t, rays[3])."""
        prm = _gi_params(params)
        out = (C.c_double * 4)()
        face = C.c_uint8()
        t = C.c_float()
        rays = (C.c_uint64 * 3)()
        lib().orc_gi_sample(C.byref(self.cstruct()), _p(prm), _p(_f4(btab)), _p(_f4(palette)),
                            _f3(o), _f3(d), int(probe), int(k), out, C.byref(face), C.byref(t),
                            rays)
        return np.array(out[:]), int(face.value), float(t.value), [int(r) for r in rays]

    def gi_update(self, params, dirs, btab, palette, probes: np.ndarray, threads=0):
        """One probe update in place on float64 records [total, 48]; returns rays[3]."""
        prm = _gi_params(params)
        assert probes.dtype == np.float64 and probes.flags["C_CONTIGUOUS"]
        rays = (C.c_uint64 * 3)()
        lib().orc_gi_update(C.byref(self.cstruct()), _p(prm), _p(_f4(dirs)), _p(_f4(btab)),
                            _p(_f4(palette)), _p(probes), _threads(threads), rays)
        return [int(r) for r in rays]

    def gi_update_learned(self, params, dirs, btab, palette, cache: "NrcOracle",
                          probes: np.ndarray, threads=0):
        """One learned probe update in place (bounce tails from `cache`); returns rays[3]."""
        prm = _gi_params(params)
        assert probes.dtype == np.float64 and probes.flags["C_CONTIGUOUS"]
        rays = (C.c_uint64 * 3)()
        ev = cache.eval_struct()
        lib().orc_gi_update_learned(C.byref(self.cstruct()), _p(prm), _p(_f4(dirs)),
                                    _p(_f4(btab)), _p(_f4(palette)), C.addressof(ev),
                                    _p(probes), _threads(threads), rays)
        return [int(r) for r in rays]

    def gi_bounce_tails(self, params, btab, palette, queries, bidx, threads=0) -> np.ndarray:
        """Traced bounce tail (rgb, float64 [n, 3]) at each cache query point."""
        prm = _gi_params(params)
        q = np.ascontiguousarray(queries, NRC_QUERY_DTYPE).reshape(-1)
        b = np.ascontiguousarray(bidx, np.uint32).reshape(-1)
        assert len(b) == len(q)
        out = np.zeros((len(q), 3), np.float64)
        lib().orc_gi_bounce_tails(C.byref(self.cstruct()), _p(prm), _p(_f4(btab)),
                                  _p(_f4(palette)), len(q), _p(q), _p(b), _p(out),
                                  _threads(threads))
        return out

    def gi_reference(self, params, dirs, btab, palette, slots, origins, threads=0):
        """FP64 estimate of SH-L1 (r, g, b, dist, dist2) x 4 per listed probe slot from the
        given origins over the full direction set `dirs`: float64 [n, 20]."""
        prm = _gi_params(params)
        dirs = _f4(dirs)
        slots = np.ascontiguousarray(slots, np.int64)
        origins = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        out = np.zeros((len(slots), 20), np.float64)
        lib().orc_gi_reference(C.byref(self.cstruct()), _p(prm), len(dirs), _p(dirs),
                               _p(_f4(btab)), _p(_f4(palette)), _p(slots), _p(origins),
                               len(slots), _p(out), _threads(threads))
        return out


def _gi_params(params) -> np.ndarray:
    a = np.frombuffer(bytes(params), dtype=GI_PARAMS_DTYPE) if not isinstance(params, np.ndarray) \
        else params.view(GI_PARAMS_DTYPE).reshape(-1)
    assert a.size == 1
    return np.ascontiguousarray(a)


def _f4(a) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def gi_sky(params, d):
    out = (C.c_double * 3)()
    lib().orc_gi_sky(_p(_gi_params(params)), _f3(d), out)
    return np.array(out[:])


def gi_bounce_index(seed, rnd, probe, k):
    return int(lib().orc_gi_bounce_index(int(seed) & 0xFFFFFFFF, int(rnd), int(probe), int(k)))


def _probe_floats(probes) -> np.ndarray:
    a = np.ascontiguousarray(probes)
    if a.dtype != np.float32:
        a = np.ascontiguousarray(a.view(np.float32))
    return a.reshape(-1, 48)


def sh_irradiance(record, n) -> np.ndarray:
    """Deringed SH-L1 irradiance / pi of one probe record (48 floats) toward normal n."""
    rec = _probe_floats(record)
    out = np.zeros(3, np.float32)
    lib().orc_sh_irradiance(_p(rec), _f3(n), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def sample_probes(probes, grid, spacing, pw, n, v) -> np.ndarray:
    """sample_probes at world point pw with normal n and view direction v (fp32)."""
    rec = _probe_floats(probes)
    out = np.zeros(3, np.float32)
    lib().orc_sample_probes(_p(rec), _i3(grid), int(spacing), _f3(pw), _f3(n), _f3(v),
                            out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def frame_indirect(probes, grid, spacing, camera, width, height, hits: np.ndarray) -> np.ndarray:
    """indirect_light in probe mode for every pixel of a frame's 20-byte hit records:
    float32 [height * width, 4]."""
    rec = _probe_floats(probes)
    cam = np.ascontiguousarray(camera, np.float32)
    h = np.ascontiguousarray(hits).view(np.uint8)
    assert h.size == 20 * width * height
    out = np.zeros((width * height, 4), np.float32)
    lib().orc_frame_indirect(_p(rec), _i3(grid), int(spacing), _p(cam), int(width), int(height),
                             _p(h), _p(out))
    return out


def water_rays(camera, width, height, hits: np.ndarray) -> np.ndarray:
    """orc_water_rays: two rays per pixel (reflection, refraction) as RAY_DTYPE records."""
    cam = np.ascontiguousarray(camera, np.float32)
    h = np.ascontiguousarray(hits).view(np.uint8)
    assert h.size == 20 * width * height
    out = np.zeros(2 * width * height, RAY_DTYPE)
    lib().orc_water_rays(_p(cam), int(width), int(height), _p(h), _p(out))
    return out


def make_rays(o, d, tmax=np.inf, flags=0) -> np.ndarray:
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    r = np.zeros(len(o), RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["tmax"] = tmax
    r["flags"] = flags
    return r


def primary_dirs(camera, width, height) -> np.ndarray:
    rays = np.zeros(width * height, RAY_DTYPE)
    cam = np.ascontiguousarray(camera, np.float32)
    lib().orc_primary_rays(_p(cam), width, height, np.inf, 0, _p(rays))
    return rays["d"].copy()


def algorithmic_bytes(it_counts: dict) -> int:
    return int(sum(it_counts[n] * b for n, b in zip(IT_NAMES, IT_BYTES)))


# ------------------------------------------------- parity against R0 (checker) --
def normalize_or_zero_v(d) -> np.ndarray:
    """glam's normalize_or_zero over rows of fp32 directions: d * (1 / sqrt(x x + y y + z z)),
    zero where that reciprocal is not finite and positive (src/raycast.rs:24)."""
    d = np.asarray(d, np.float32).reshape(-1, 3)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ln = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        rcp = np.float32(1.0) / ln
    ok = np.isfinite(rcp) & (rcp > 0)
    return np.where(ok[:, None], d * rcp[:, None], np.float32(0.0)).astype(np.float32)


def ray_box_entry_v(o, d, v) -> np.ndarray:
    """src/accel.rs:578-586 in fp32, one row per ray: the entry t of voxel v along o + t d."""
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    vmin = np.asarray(v).reshape(-1, 3).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        inv = np.where(np.abs(d) < 1e-8, np.float32(1e30), np.float32(1.0) / d).astype(np.float32)
        t0 = (vmin - o) * inv
        t1 = (vmin + np.float32(1.0) - o) * inv
    return np.maximum(np.max(np.minimum(t0, t1), axis=1), np.float32(0.0)).astype(np.float32)


def lattice_touch_v(p, eps=0.02) -> np.ndarray:
    """src/accel.rs:588-599 per row: how many coordinates of p lie on a voxel-lattice plane."""
    p = np.asarray(p, np.float64).reshape(-1, 3)
    return np.sum(np.abs(p - np.round(p)) < eps, axis=1)


def classify_vs_r0(eyes, dirs, cand: np.ndarray, ref: np.ndarray, tmax=None) -> dict:
    """Hits of a candidate traversal against the reference CPU traversal R0 (orc_raycast_r0,
    src/raycast.rs:21-141) by the rule of the reference's accel_matches_cpu_raycast test
    (src/accel.rs:655-718), for every ray at once.

    eyes: ray origins (one, or one per ray); dirs: the rays' unit directions (after
    normalize_or_zero); cand, ref: HIT_DTYPE records. A hit point is eye + t dir, with R0's t
    the entry t of its voxel (ray_box_entry). It is grazing when two of its coordinates lie
    within 0.02 of a lattice plane (lattice_touch >= 2): an edge or corner, where the
    accumulated t of R0 and an exact walk may order two plane crossings differently.
      agree           both hit the same voxel (clean, or a tie when R0's point grazes)
      ties            agreeing grazing hits + disagreements where either point grazes
      clean_disagree  other disagreements (one side hits, the other misses, or two voxels)
      tmax_ties       disagreements where exactly one side hits within 1e-3 (relative, at
                      least 1e-3 absolute) of the ray's tmax: R0 compares its accumulated t
                      against tmax, an exact walk its exact t; counted apart, not clean
    On clean agreements the face and material must match and |t - t_R0| is measured, t_R0
    being the t R0 reports."""
    dirs = np.asarray(dirs, np.float32).reshape(-1, 3)
    n = len(dirs)
    eyes = np.broadcast_to(np.asarray(eyes, np.float32).reshape(-1, 3), (n, 3))
    ch = cand["hit"] != 0
    rh = ref["hit"] != 0
    t_ref = ray_box_entry_v(eyes, dirs, ref["voxel"])
    cpu_p = eyes + t_ref[:, None] * dirs
    cand_p = eyes + cand["t"].astype(np.float32)[:, None] * dirs
    cpu_graze = rh & (lattice_touch_v(cpu_p) >= 2)
    cand_graze = ch & (lattice_touch_v(cand_p) >= 2)
    same = ch & rh & np.all(cand["voxel"] == ref["voxel"], axis=1)
    both_miss = ~ch & ~rh
    agree_clean = same & ~cpu_graze
    agree_tie = same & cpu_graze
    dis = ~both_miss & ~same
    dis_tie = dis & (cpu_graze | cand_graze)
    dis_clean = dis & ~(cpu_graze | cand_graze)
    tmax_tie = np.zeros(n, bool)
    if tmax is not None:
        tm = np.broadcast_to(np.asarray(tmax, np.float64).reshape(-1), (n,))
        tol = np.maximum(1e-3, 1e-3 * np.abs(tm))
        one_side = ch ^ rh
        t_side = np.where(ch, cand["t"].astype(np.float64), t_ref.astype(np.float64))
        tmax_tie = dis_clean & one_side & (np.abs(t_side - tm) <= tol)
        dis_clean = dis_clean & ~tmax_tie
    # against the t R0 reports: its ray_box_entry along its own normalize_or_zero of the
    # direction, which may differ from the given unit direction in the last bit
    dt = np.abs(cand["t"].astype(np.float64) - ref["t"].astype(np.float64))
    face_bad = agree_clean & (cand["face"] != ref["face"])
    mat_bad = agree_clean & (cand["material"] != ref["material"])
    return {
        "rays": int(n), "both_miss": int(both_miss.sum()),
        "agree": int(same.sum()), "agree_clean": int(agree_clean.sum()),
        "ties": int(agree_tie.sum() + dis_tie.sum()), "agree_ties": int(agree_tie.sum()),
        "disagree_ties": int(dis_tie.sum()), "tmax_ties": int(tmax_tie.sum()),
        "clean_disagree": int(dis_clean.sum()),
        "face_mismatch": int(face_bad.sum()), "material_mismatch": int(mat_bad.sum()),
        "max_dt": float(dt[agree_clean].max()) if agree_clean.any() else 0.0,
        "max_dt_all_agree": float(dt[same].max()) if same.any() else 0.0,
        "bad_index": np.flatnonzero(dis_clean | face_bad | mat_bad)[:16].tolist(),
    }


def shadow_origins(eye, dirs, hits: np.ndarray) -> np.ndarray:
    """orc_shadow_origin per hit record: p_hit + n * 0.001 (shaders/raymarch.wgsl:3955)."""
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = len(dirs)
    eyes = np.ascontiguousarray(np.broadcast_to(np.asarray(eye, np.float32).reshape(-1, 3),
                                                (n, 3)))
    hits = np.ascontiguousarray(hits, HIT_DTYPE)
    out = np.zeros((n, 3), np.float32)
    F32P = C.POINTER(C.c_float)
    L = lib()
    base_e, base_d, base_o = eyes.ctypes.data, dirs.ctypes.data, out.ctypes.data
    for i in range(n):
        L.orc_shadow_origin(C.cast(base_e + 12 * i, F32P), C.cast(base_d + 12 * i, F32P),
                            hits.ctypes.data + 20 * i, C.cast(base_o + 12 * i, F32P))
    return out


def classify_shadows_vs_r0(ow: "OracleWorld", eye, dirs, prim: np.ndarray, occ, occ_r0,
                           sun, shadow_tmax, rows_mask) -> dict:
    """Shadow bytes of a candidate frame against R0's shadow bytes, on the rays of rows_mask
    whose primary hits agree with R0 on a clean hit with the same t (so both traced the same
    shadow ray). A disagreeing shadow ray is traced again by R0 and by the exact walk (any-hit)
    and classified like a primary hit (classify_vs_r0)."""
    occ = np.asarray(occ).reshape(-1).astype(bool)
    occ_r0 = np.asarray(occ_r0).reshape(-1).astype(bool)
    idx = np.flatnonzero(rows_mask & (prim["hit"] != 0))
    dis = idx[occ[idx] != occ_r0[idx]]
    out = {"rays": int(len(idx)), "agree": int(len(idx) - len(dis)), "ties": 0,
           "clean_disagree": 0}
    if len(dis):
        so = shadow_origins(eye, np.asarray(dirs, np.float32).reshape(-1, 3)[dis], prim[dis])
        sd = normalize_or_zero_v(np.asarray(sun, np.float32).reshape(1, 3))[0]
        a = np.zeros(len(dis), HIT_DTYPE)
        b = np.zeros(len(dis), HIT_DTYPE)
        for k in range(len(dis)):
            a[k], _ = ow.trace_r1(so[k], sd, tmax=shadow_tmax, flags=ANY_HIT)
            b[k] = ow.raycast_r0(so[k], sd, tmax=shadow_tmax)
        c = classify_vs_r0(so, np.broadcast_to(sd, so.shape), a, b, tmax=shadow_tmax)
        # a re-traced pair that agrees (or misses on both sides) leaves the frame's
        # disagreement unexplained: only traversal ties explain one
        out["ties"] = c["disagree_ties"] + c["tmax_ties"]
        out["clean_disagree"] = int(len(dis)) - out["ties"]
    return out


# noise / generator helpers
def hash2(x, z):
    return lib().orc_hash2(int(x), int(z))


def hash3(x, y, z):
    return lib().orc_hash3(int(x), int(y), int(z))


def value_noise_2d(x, z):
    return lib().orc_value_noise_2d(float(x), float(z))


def value_noise_3d(x, y, z):
    return lib().orc_value_noise_3d(float(x), float(y), float(z))


def fbm_2d(x, z, octaves):
    return lib().orc_fbm_2d(float(x), float(z), int(octaves))


def seed_offset_xz(seed):
    out = (C.c_float * 2)()
    lib().orc_seed_offset_xz(C.c_uint64(seed), out)
    return float(out[0]), float(out[1])


def terrain_height(dims, seed, x, z):
    return int(lib().orc_terrain_height(_i3(dims), C.c_uint64(seed), int(x), int(z)))


def brick_uniform_of(brick: np.ndarray) -> int:
    b = np.ascontiguousarray(np.asarray(brick, BRICK_DTYPE).reshape(1))
    return int(lib().orc_brick_uniform_of(_p(b)))


class NrcOracle:
    """FP64 radiance cache (oracle_nrc.c) over a config and logical fp32 parameters.
    emulate=True rounds what the device holds in bf16 (parameters, layer inputs, layer
    gradients) at the same points."""

    def __init__(self, config: np.ndarray, mlp: np.ndarray, grid: np.ndarray, emulate=False):
        self.cfg = np.ascontiguousarray(config.view(NRC_CONFIG_DTYPE).reshape(1))
        c = self.cfg[0]
        self.grid_params = int(c["levels"]) * (1 << int(c["log2_table"])) * 2
        self.mlp = np.ascontiguousarray(mlp, np.float32).reshape(NRC_MLP_PARAMS)
        self.grid = np.ascontiguousarray(grid, np.float32).reshape(self.grid_params)
        self.emulate = int(bool(emulate))

    @staticmethod
    def _q(queries):
        q = np.ascontiguousarray(queries, NRC_QUERY_DTYPE).reshape(-1)
        return q

    def encode(self, query) -> np.ndarray:
        q = self._q(query)
        out = np.zeros(NRC_WIDTH, np.float64)
        lib().orc_nrc_encode(_p(self.cfg), _p(self.grid), _p(q[:1]), self.emulate, _p(out))
        return out

    def forward(self, queries) -> np.ndarray:
        q = self._q(queries)
        out = np.zeros((len(q), 3), np.float64)
        lib().orc_nrc_forward(_p(self.cfg), _p(self.mlp), _p(self.grid), len(q), _p(q),
                              self.emulate, _p(out))
        return out

    def loss(self, queries, targets) -> float:
        q = self._q(queries)
        t = np.ascontiguousarray(targets, np.float32).reshape(len(q), 4)
        return lib().orc_nrc_loss(_p(self.cfg), _p(self.mlp), _p(self.grid), len(q), _p(q),
                                  _p(t), self.emulate)

    def loss_grad(self, queries, targets):
        """(loss, d loss / d mlp [13312], d loss / d grid [levels * 2^log2_table * 2])."""
        q = self._q(queries)
        t = np.ascontiguousarray(targets, np.float32).reshape(len(q), 4)
        gm = np.zeros(NRC_MLP_PARAMS, np.float64)
        gg = np.zeros(self.grid_params, np.float64)
        loss = lib().orc_nrc_loss_grad(_p(self.cfg), _p(self.mlp), _p(self.grid), len(q), _p(q),
                                       _p(t), self.emulate, _p(gm), _p(gg))
        return loss, gm, gg

    def eval_struct(self):
        """orc_nrc_eval for orc_gi_update_learned (keeps this object alive while used)."""
        class _Eval(C.Structure):
            _fields_ = [("cfg", C.c_void_p), ("mlp", C.c_void_p), ("grid", C.c_void_p),
                        ("emulate", C.c_int)]
        return _Eval(_p(self.cfg), _p(self.mlp), _p(self.grid), self.emulate)


def nrc_scales(levels, coarse_cell, fine_cell) -> np.ndarray:
    out = np.zeros(int(levels), np.float32)
    lib().orc_nrc_scales(int(levels), float(coarse_cell), float(fine_cell), _p(out))
    return out


def nrc_hash(x, y, z, log2_table) -> int:
    return int(lib().orc_nrc_hash(int(x), int(y), int(z), int(log2_table)))


def bf16_round(v) -> float:
    return float(lib().orc_bf16_round(float(v)))


def nrc_adam(p, m, v, g, lr, beta1, beta2, eps, step, sparse):
    """One Adam step in place on float64 arrays."""
    for a in (p, m, v, g):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    lib().orc_nrc_adam(len(p), _p(p), _p(m), _p(v), _p(g), float(lr), float(beta1), float(beta2),
                       float(eps), int(step), int(bool(sparse)))
