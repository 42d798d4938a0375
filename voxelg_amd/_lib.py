"""ctypes binding of libvoxelg_amd.so (include/vg.h).

The library is the product: the HIP kernels for gfx950 and the host-side world builders.
Loading fails loudly when it has not been built; nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VOXELG_AMD_LIB points the binding at another build of the same library (a variant built
# elsewhere, for instance); the default is the in-tree build.
LIB_PATH = os.environ.get("VOXELG_AMD_LIB") or os.path.join(_HERE, "libvoxelg_amd.so")

BRICK_DTYPE = np.dtype([("occupancy", "<u8"), ("materials", "u1", (64,))])
RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("tmax", "<f4"), ("d", "<f4", (3,)), ("flags", "<u4")])
HIT_DTYPE = np.dtype(
    [("voxel", "<i4", (3,)), ("t", "<f4"), ("face", "u1"), ("material", "u1"), ("hit", "u1"),
     ("pad", "u1")]
)
BATCH_DTYPE = np.dtype([("rays", "<u8"), ("hits", "<u8"), ("n", "<u4"), ("flags", "<u4")])
# vg_probe: GiProbe (shaders/gi_probes.wgsl:25-48), 12 x vec4 f32.
PROBE_DTYPE = np.dtype([
    ("sh_r", "<f4", (4,)), ("sh_g", "<f4", (4,)), ("sh_b", "<f4", (4,)),
    ("sh_dist", "<f4", (4,)), ("sh_dist2", "<f4", (4,)), ("info", "<f4", (4,)),
    ("st_r", "<f4", (4,)), ("st_g", "<f4", (4,)), ("st_b", "<f4", (4,)),
    ("st_dist", "<f4", (4,)), ("st_dist2", "<f4", (4,)), ("st_info", "<f4", (4,))])
# vg_gi_params
GI_PARAMS_DTYPE = np.dtype([
    ("grid", "<i4", (3,)), ("spacing", "<i4"), ("rays_per_probe", "<i4"), ("epochs", "<i4"),
    ("update_div", "<i4"), ("round", "<i4"), ("bounces", "<i4"), ("seed", "<u4"),
    ("flags", "<u4"), ("hysteresis", "<f4"), ("gi_dist", "<f4"), ("shadow_tmax", "<f4"),
    ("sun", "<f4", (3,)), ("gi_sky", "<f4"), ("sun_color", "<f4", (3,)),
    ("ambient_tail", "<f4"), ("ambient", "<f4", (3,)), ("day_t", "<f4"),
    ("sky_horizon", "<f4", (3,)), ("dusk_w", "<f4"), ("sky_zenith", "<f4", (3,)),
    ("halo", "<f4"), ("sun_xz", "<f4", (2,)), ("pad", "<f4", (2,))])
# vg_nrc_config / vg_nrc_query (radiance cache)
NRC_CONFIG_DTYPE = np.dtype([
    ("levels", "<i4"), ("log2_table", "<i4"), ("scale", "<f4", (16,)), ("lr_grid", "<f4"),
    ("lr_mlp", "<f4"), ("beta1", "<f4"), ("beta2", "<f4"), ("eps", "<f4"), ("init_grid", "<f4"),
    ("seed", "<u4"), ("batch", "<i4"), ("steps", "<i4"), ("pad", "<i4")])
NRC_QUERY_DTYPE = np.dtype([("pos", "<f4", (3,)), ("face", "<u4"), ("dir", "<f4", (3,)),
                            ("aux", "<u4")])
assert NRC_CONFIG_DTYPE.itemsize == 112 and NRC_QUERY_DTYPE.itemsize == 32
# vg_edit: a sphere edit (apply_sphere, src/app.rs:199-210)
EDIT_DTYPE = np.dtype([("center", "<i4", (3,)), ("radius", "u1"), ("material", "u1"),
                       ("pad", "u1", (2,))])
assert EDIT_DTYPE.itemsize == 16
VG_NRC_MLP_PARAMS = 3 * 64 * 64 + 16 * 64
VG_NRC_MAX_BATCH = 1 << 20
VG_NRC_TRAIN = 1
assert BRICK_DTYPE.itemsize == 72 and RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 20
assert BATCH_DTYPE.itemsize == 24 and PROBE_DTYPE.itemsize == 192
assert GI_PARAMS_DTYPE.itemsize == 152
VG_GI_SKIP_BURIED = 1
VG_GI_BOUNCE_TABLE = 4096

VG_OK = 0
VG_SCHEDULE_PERSISTENT = 0
VG_SCHEDULE_STATIC = 1
VG_OPT_SCHEDULE = 1
VG_OPT_FRAME_WAVES_PER_CU = 2
VG_OPT_FRAME_REFILL = 3
VG_OPT_RAYS_WAVES_PER_CU = 4
VG_OPT_RAYS_REFILL = 5
VG_OPT_CU_COUNT = 6
VG_OPT_PROBE_WAVES_PER_CU = 7
VG_OPT_PROBE_REFILL = 8
VG_OPT_PROBE_CHUNK = 9
VG_OPT_FRAME_STRIPES = 10
VG_OPT_FRAME_BLOCK_W = 11
VG_OPT_PROBE_TBATCH = 12
VG_OPT_PROBE_MATSPEC = 13
VG_RAY_LOD = 1
VG_RAY_SKIP_TRANSPARENT = 2
VG_RAY_ANY_HIT = 4
VG_RAY_BEAM = 8
VG_FACE_NONE = 6
VG_IT_COUNT = 14
IT_NAMES = ("l4_skip", "chunk_skip", "tile_skip", "tile_lod", "tile_uniform", "brick_skip",
            "brick_uniform", "brick_lod", "voxel", "chunk_group_skip", "tile_group_skip",
            "brick_group_skip", "voxel_group_skip", "beam_test")
# Algorithmic bytes per iteration class: the pyramid words the walk's test chain reads in
# the reference's order (shaders/raymarch.wgsl:4626-4775): L4 word 8, chunk word 8,
# tile-uniform byte 1, tile word 8, brick-uniform byte 1, occupancy word 8. A tile-LOD test
# adds its tile word, the first child brick's occupancy and one material byte (17); a brick
# LOD reads the occupancy and one material byte past the brick-uniform byte (35). A 2x2x2
# group skip reads the same word as the single-cell skip of its level.
IT_BYTES = (8, 8, 16, 17, 17, 25, 26, 35, 34, 8, 16, 25, 34, 8)
HIT_RECORD_BYTES = 20   # one vg_hit written per ray
MATERIAL_BYTES = 1      # material byte read at a voxel hit
RAY_INPUT_BYTES = 32    # one vg_ray read per batched ray


def algorithmic_bytes(it_counts: dict) -> int:
    return int(sum(int(it_counts[n]) * b for n, b in zip(IT_NAMES, IT_BYTES)))


class VGError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


_lib = None
_lock = threading.Lock()

_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_U64P = C.POINTER(C.c_uint64)
_F32P = C.POINTER(C.c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "vg_abi_version": (C.c_int, []),
    "vg_last_error": (C.c_char_p, []),
    "vg_status_string": (C.c_char_p, [C.c_int]),
    "vg_world_counts": (C.c_int, [_I32P, _I64P]),
    "vg_generate_terrain": (C.c_int, [_I32P, _I32P, C.c_uint64, _P, C.c_int]),
    "vg_terrain_height": (C.c_int, [_I32P, C.c_uint64, C.c_int32, C.c_int32, _I32P]),
    "vg_build_pyramid": (C.c_int, [_I32P, _P, _P, _P, _P, _P, _P, C.c_int]),
    "vg_camera_from_pose": (C.c_int, [_F32P, C.c_float, C.c_float, C.c_float, C.c_int32,
                                      C.c_int32, _I32P, _P]),
    "vg_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "vg_ctx_destroy": (C.c_int, [_P]),
    "vg_world_create": (C.c_int, [_P, _I32P, C.POINTER(_P)]),
    "vg_world_destroy": (C.c_int, [_P]),
    "vg_world_upload": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _I32P]),
    "vg_world_generate": (C.c_int, [_P, C.c_uint64, _I32P, _P]),
    "vg_trace_rays": (C.c_int, [_P, _P, C.c_size_t, C.c_uint32, _P, _P]),
    "vg_trace_primary": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_uint32, _P, _P]),
    "vg_trace_shadow": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _P, _F32P, C.c_float, _P, _P]),
    "vg_trace_frame": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_uint32, _F32P,
                                 C.c_float, _P, _P, _P]),
    "vg_trace_frame_noalloc": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_uint32,
                                         _F32P, C.c_float, _P, _P, _P]),
    "vg_trace_frame_counts": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_uint32,
                                        _F32P, C.c_float, _P, _P, _U64P]),
    "vg_trace_rays_counts": (C.c_int, [_P, _P, C.c_size_t, C.c_uint32, _P, _U64P]),
    "vg_raycast": (C.c_int, [_P, _F32P, _F32P, _P]),
    "vg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "vg_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "vg_free": (C.c_int, [_P, _P]),
    "vg_copy_to_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "vg_copy_to_host": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "vg_memset": (C.c_int, [_P, _P, C.c_int, C.c_size_t]),
    "vg_copy_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "vg_synchronize": (C.c_int, [_P]),
    "vg_stream_create": (C.c_int, [_P, C.POINTER(_P)]),
    "vg_stream_destroy": (C.c_int, [_P, _P]),
    "vg_stream_synchronize": (C.c_int, [_P, _P]),
    "vg_event_create": (C.c_int, [_P, C.POINTER(_P)]),
    "vg_event_destroy": (C.c_int, [_P, _P]),
    "vg_event_record": (C.c_int, [_P, _P, _P]),
    "vg_event_elapsed_ms": (C.c_int, [_P, _P, _P, C.POINTER(C.c_float)]),
    "vg_ctx_set_option": (C.c_int, [_P, C.c_int, C.c_int]),
    "vg_ctx_get_option": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "vg_autotune_frame": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_uint32, _F32P,
                                    C.c_float, _P, _P, C.POINTER(C.c_float)]),
    "vg_autotune_rays": (C.c_int, [_P, _P, C.c_size_t, C.c_uint32, _P, C.POINTER(C.c_float)]),
    "vg_trace_batches": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, _P]),
    "vg_trace_batches_counts": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, _U64P]),
    "vg_autotune_batches": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "vg_graph_begin": (C.c_int, [_P, _P]),
    "vg_graph_end": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "vg_graph_launch": (C.c_int, [_P, _P]),
    "vg_graph_destroy": (C.c_int, [_P]),
    "vg_gi_params_default": (C.c_int, [_F32P, _P]),
    "vg_gi_default_palette": (C.c_int, [_P]),
    "vg_gi_directions": (C.c_int, [C.c_int32, _P]),
    "vg_gi_bounce_table": (C.c_int, [_P]),
    "vg_probes_create": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "vg_probes_destroy": (C.c_int, [_P]),
    "vg_probes_set_palette": (C.c_int, [_P, _P]),
    "vg_probes_reset": (C.c_int, [_P, _P]),
    "vg_probes_data": (C.c_int, [_P, C.POINTER(_P)]),
    "vg_probes_download": (C.c_int, [_P, _P]),
    "vg_probes_upload": (C.c_int, [_P, _P]),
    "vg_probes_update": (C.c_int, [_P, _P, _P]),
    "vg_trace_water_secondary": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _P, _P, _P, _P]),
    "vg_frame_indirect": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, _P, _P, _P]),
    "vg_probes_update_counts": (C.c_int, [_P, _P, _U64P, _U64P]),
    "vg_nrc_config_default": (C.c_int, [_P]),
    "vg_nrc_scales": (C.c_int, [C.c_int32, C.c_float, C.c_float, _P]),
    "vg_nrc_create": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "vg_nrc_destroy": (C.c_int, [_P]),
    "vg_nrc_param_counts": (C.c_int, [_P, _I64P]),
    "vg_nrc_get_params": (C.c_int, [_P, _P, _P]),
    "vg_nrc_set_params": (C.c_int, [_P, _P, _P]),
    "vg_nrc_infer": (C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    "vg_nrc_train_step": (C.c_int, [_P, _P, _P, C.c_uint32, _P]),
    "vg_nrc_join": (C.c_int, [_P, _P]),
    "vg_nrc_gradients": (C.c_int, [_P, _P, _P, C.c_uint32, _P, _P, _P]),
    "vg_nrc_stats": (C.c_int, [_P, _I64P, _F32P]),
    "vg_nrc_last_batch": (C.c_int, [_P, _P, _P]),
    "vg_probes_update_learned": (C.c_int, [_P, _P, _P, C.c_uint32, _P]),
    "vg_probes_update_learned_counts": (C.c_int, [_P, _P, _P, C.c_uint32, _U64P, _U64P,
                                                  _U64P]),
    "vg_probes_set_timing": (C.c_int, [_P, C.c_int]),
    "vg_probes_timing": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "vg_autotune_probes": (C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(C.c_float)]),
    "vg_world_apply_edits": (C.c_int, [_P, _P, C.c_uint32, _P]),
    "vg_frame_mask_words": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_uint32)]),
    "vg_frame_dirty_blocks": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _F32P, C.c_float, _P, _P]),
    "vg_trace_frame_blocks": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_uint32,
                                        _F32P, C.c_float, _P, _P, _P, _P]),
    "vg_world_download": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
}


def lib():
    """Load libvoxelg_amd.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                    "g.build()'` (or make -C voxelg_amd/csrc)"
                )
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(status: int, where: str) -> None:
    if status != VG_OK:
        msg = lib().vg_last_error().decode(errors="replace")
        raise VGError(status, where, msg)


def i32x3(v) -> C.Array:
    return (C.c_int32 * 3)(*[int(x) for x in v])


def f32x3(v) -> C.Array:
    return (C.c_float * 3)(*[float(np.float32(x)) for x in v])


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
