/*
 * vg.h — C ABI of libvoxelg_amd.so, the MI355X (gfx950) voxel ray-query engine.
 *
 * The reference (marcstampfli1/voxelG) has no FFI layer; its ray query sits behind three
 * de-facto boundaries, each replaced here by plain-pointer entry points:
 *   1. the CPU ray query      pub fn raycast(camera_world, dir, &World, world_origin)
 *                             -> Option<PickHit>                       src/raycast.rs:21
 *   2. the batched GPU query  Ray[] -> Hit[] storage buffers           shaders/accel_probe.wgsl:21-126
 *                             (host side probe_with_accel,            src/accel.rs:417-513)
 *   3. the frame kernels      cs_beam + cs_main's trace + trace_any    shaders/beam.wgsl:31-113,
 *                             over the uploaded world                 shaders/raymarch.wgsl:576-762
 *                             (Renderer::upload_world / render,       src/renderer.rs:1248, 1533)
 * plus the host-side world builders the upload path feeds on (src/voxel.rs:233-272, 950-1000).
 *
 * Conventions
 *   - Every entry point returns an int status (VG_OK or a negative VG_E_*); nothing throws
 *     across the ABI. vg_last_error() returns a thread-local message for the last failure.
 *   - Device pointers (rays, hits, occlusion bytes) are plain device addresses the caller
 *     allocated on the world's device; host pointers are marked _host.
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream. Kernels are
 *     enqueued asynchronously; launches on one world must not run concurrently with each
 *     other or with vg_world_upload (they share the world's scratch counter).
 *   - World dimensions are voxels per axis, each a multiple of 64; an axis whose window
 *     origin is not a multiple of its dimension (toroidal fold) must be a multiple of 256.
 *   - Distinct worlds may be traced from distinct host threads; creation, upload and
 *     destruction are not thread-safe against tracing on the same world.
 */
#ifndef VG_H
#define VG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VG_ABI_VERSION 1

enum {
    VG_OK = 0,
    VG_E_INVALID = -1,     /* bad argument (null pointer, bad dims, size mismatch) */
    VG_E_NOMEM = -2,       /* host or device allocation failed */
    VG_E_DEVICE = -3,      /* HIP runtime / launch error */
    VG_E_UNSUPPORTED = -4, /* valid request the library does not serve */
    VG_E_STATE = -5        /* world not uploaded, or wrong context */
};

/* Ray flags (vg_ray.flags and the per-call flags are OR-ed). */
#define VG_RAY_LOD 1u              /* brick LOD past 400, tile LOD past 520 (raymarch.wgsl:4489-4494) */
#define VG_RAY_SKIP_TRANSPARENT 2u /* pass water and glass (trace_no_water, raymarch.wgsl:3450) */
#define VG_RAY_ANY_HIT 4u          /* occlusion semantics (trace_any, raymarch.wgsl:4345) */
#define VG_RAY_BEAM 8u             /* frame primary rays: fused conservative beam prologue per
                                      8x8 pixel block (cs_beam, shaders/beam.wgsl:31-113); it
                                      never changes a result. Ignored by vg_trace_rays. */

/* Face codes (encode_face_normal, shaders/raymarch.wgsl:482-486): the hit face's outward
 * normal is +x 0, -x 1, +y 2, -y 3, +z 4, -z 5; 6 = no face (miss, or no entry plane). */
#define VG_FACE_NONE 6

/* voxelG's Brick (src/voxel.rs:296-300): 72 bytes, voxel bit / material index x + 4z + 16y. */
typedef struct vg_brick {
    uint64_t occupancy;
    uint8_t materials[64];
} vg_brick;

/* A ray of the batched query: accel_probe's Ray (shaders/accel_probe.wgsl:21-26) with the
 * two pad words carrying a per-ray tmax and flags. d need not be normalized (it is
 * normalized like glam's normalize_or_zero; a zero d misses). World coordinates. */
typedef struct vg_ray {
    float o[3];
    float tmax;
    float d[3];
    uint32_t flags;
} vg_ray;

/* A hit record (accel_probe's Hit, shaders/accel_probe.wgsl:28-39, packed to 20 bytes):
 * world voxel, entry-t of that voxel, face code, material, hit flag. A miss has hit 0,
 * t -1, face VG_FACE_NONE and zeros elsewhere. */
typedef struct vg_hit {
    int32_t voxel[3];
    float t;
    uint8_t face;
    uint8_t material;
    uint8_t hit;
    uint8_t pad;
} vg_hit;

/* CameraUniform (src/camera.rs:61-113), 176 bytes, same field order. The traversal reads
 * origin, forward, right, up, tan_half_fov, resolution and jitter. */
typedef struct vg_camera {
    float origin[3];
    float _pad0;
    float forward[3];
    float _pad1;
    float right[3];
    float _pad2;
    float up[3];
    float tan_half_fov;
    float resolution[2];
    float time;
    float wind_x;
    int32_t world_origin[3];
    int32_t gi_round;
    float jitter[2];
    float taa_blend;
    float reproject_lighting;
    float prev_origin[3];
    float wind_z;
    float prev_forward[3];
    float sun_time;
    float prev_right[3];
    float prev_valid;
    float prev_up[3];
    float _pad9;
} vg_camera;

/* Iteration classes of the hierarchical walk, reported by the *_counts entry points. */
enum {
    VG_IT_L4_SKIP = 0,
    VG_IT_CHUNK_SKIP,
    VG_IT_TILE_SKIP,
    VG_IT_TILE_LOD,
    VG_IT_TILE_UNIFORM,
    VG_IT_BRICK_SKIP,
    VG_IT_BRICK_UNIFORM,
    VG_IT_BRICK_LOD,
    VG_IT_VOXEL,
    VG_IT_CHUNK_GROUP_SKIP, /* empty 2x2x2 chunk group (128^3) left in one step */
    VG_IT_TILE_GROUP_SKIP,  /* empty 2x2x2 tile group (32^3) */
    VG_IT_BRICK_GROUP_SKIP, /* empty 2x2x2 brick group (8^3) */
    VG_IT_VOXEL_GROUP_SKIP, /* empty 2x2x2 voxel octant of a brick (2^3) */
    VG_IT_BEAM_TEST,        /* beam prologue: one tile tested by one lane */
    VG_IT_COUNT
};

typedef struct vg_ctx vg_ctx;
typedef struct vg_world vg_world;

/* ------------------------------------------------ device memory / timing -- */
/* Plain device buffers, streams and events for callers that have no allocator of their
 * own (the engine needs nothing else from its host). Copies are synchronous. */
int vg_device_count(int *out);
int vg_alloc(vg_ctx *ctx, size_t bytes, void **dev_out);
int vg_free(vg_ctx *ctx, void *dev);
int vg_copy_to_device(vg_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int vg_copy_to_host(vg_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
int vg_memset(vg_ctx *ctx, void *dev, int value, size_t bytes);
int vg_copy_device(vg_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);
int vg_synchronize(vg_ctx *ctx);
int vg_stream_create(vg_ctx *ctx, void **stream_out); /* default priority */
int vg_stream_destroy(vg_ctx *ctx, void *stream);
int vg_stream_synchronize(vg_ctx *ctx, void *stream);
int vg_event_create(vg_ctx *ctx, void **event_out);
int vg_event_destroy(vg_ctx *ctx, void *event);
int vg_event_record(vg_ctx *ctx, void *event, void *stream);
/* milliseconds between two recorded events (waits for `end`) */
int vg_event_elapsed_ms(vg_ctx *ctx, void *start, void *end, float *ms_out);

/* ----------------------------------------------------------------- misc -- */
int vg_abi_version(void);
const char *vg_last_error(void);
const char *vg_status_string(int status);

/* ---------------------------------------------- host-side world building -- */
/* Bricks, tiles, chunks, L4 cells and voxels of a window (src/world_dims.rs:9-37). */
int vg_world_counts(const int32_t dims[3], int64_t out[5]);

/* Benchmark terrain into caller-owned bricks (counts[0] entries, reference x-fastest order).
 * World voxel origin+local is stored at its toroidal slot. Replaces World::fill_demo_terrain
 * (src/voxel.rs:1005) with the benchmark generator: fBm heightmap, 3D-noise caves,
 * grass/dirt/stone/sand, water below y=56. threads <= 0 uses all hardware threads. */
int vg_generate_terrain(const int32_t dims[3], const int32_t origin[3], uint64_t seed,
                        vg_brick *bricks_host, int threads);
/* Terrain surface height of world column (x, z) for the benchmark generator. */
int vg_terrain_height(const int32_t dims[3], uint64_t seed, int32_t x, int32_t z, int32_t *h_out);

/* World::rebuild_all_masks + rebuild_all_uniform (src/voxel.rs:950-1000) over caller-owned
 * arrays in the reference's x-fastest order. */
int vg_build_pyramid(const int32_t dims[3], const vg_brick *bricks_host, uint64_t *tile_mask,
                     uint64_t *chunk_mask, uint64_t *l4_mask, uint8_t *brick_uniform,
                     uint8_t *tile_uniform, int threads);

/* CameraUniform::from_camera (src/camera.rs:130-166) for a fly camera at pos with yaw,
 * pitch and vertical fov (radians); basis per Camera::forward/right/up (:25-42). */
int vg_camera_from_pose(const float pos[3], float yaw, float pitch, float fov_y, int32_t width,
                        int32_t height, const int32_t world_origin[3], vg_camera *out);

/* ------------------------------------------------------------- device -- */
/* A context owns one HIP device. */
int vg_ctx_create(int device, vg_ctx **out);
int vg_ctx_destroy(vg_ctx *ctx);

/* Allocate a world's device storage (pyramid, LUTs, occupancy, materials). */
int vg_world_create(vg_ctx *ctx, const int32_t dims[3], vg_world **out);
int vg_world_destroy(vg_world *world);

/* Renderer::upload_world (src/renderer.rs:1248-1395), full upload: copies the caller's
 * reference-layout arrays (counts from vg_world_counts) into the engine's own HBM layout.
 * Synchronous; the caller keeps ownership of the host arrays. */
int vg_world_upload(vg_world *world, const vg_brick *bricks, const uint64_t *tile_mask,
                    const uint64_t *chunk_mask, const uint64_t *l4_mask,
                    const uint8_t *brick_uniform, const uint8_t *tile_uniform,
                    const int32_t origin[3]);

/* Batched ray query (shaders/accel_probe.wgsl:66-126): n rays in, n hits out, both device
 * arrays. First solid voxel along each ray within its tmax, water included unless
 * VG_RAY_SKIP_TRANSPARENT. */
int vg_trace_rays(vg_world *world, const vg_ray *rays, size_t n, uint32_t flags, vg_hit *hits,
                  void *stream);

/* cs_main's primary trace (shaders/raymarch.wgsl:576-626 + trace :4600): one pinhole ray
 * per pixel from the camera, hits[y * width + x]. */
int vg_trace_primary(vg_world *world, const vg_camera *camera_host, int32_t width,
                     int32_t height, float tmax, uint32_t flags, vg_hit *hits, void *stream);

/* shade()'s sun shadow ray (shaders/raymarch.wgsl:3953-3995 + trace_any :4345): for every
 * primary hit, an any-hit ray from p_hit + n * 0.001 toward sun; occluded[i] = 0/1. */
int vg_trace_shadow(vg_world *world, const vg_camera *camera_host, int32_t width,
                    int32_t height, const vg_hit *primary_hits, const float sun[3], float tmax,
                    uint8_t *occluded, void *stream);

/* Primary + shadow for one frame in one launch. */
int vg_trace_frame(vg_world *world, const vg_camera *camera_host, int32_t width, int32_t height,
                   float primary_tmax, uint32_t primary_flags, const float sun[3],
                   float shadow_tmax, vg_hit *hits, uint8_t *occluded, void *stream);

/* vg_trace_frame for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It has no counterpart in the reference.
 * It calls no hipMalloc / hipFree: its scratch (the ray counter and the beam start t of up to
 * 512 x 512 8x8 pixel blocks, 1 MiB) is a static array per device that the library holds until
 * the process ends, its address taken once per device by vg_ctx_create. Frames above 262 144
 * blocks (4096 x 4096 pixels) return VG_E_INVALID. All calls on one device share that array,
 * which keeps the last call's contents (each call resets what must start at zero): calls that
 * overlap on one device race; serialize them on one stream. */
int vg_trace_frame_noalloc(vg_world *world, const vg_camera *camera_host, int32_t width,
                           int32_t height, float primary_tmax, uint32_t primary_flags,
                           const float sun[3], float shadow_tmax, vg_hit *hits,
                           uint8_t *occluded, void *stream);

/* The same frame, also reporting the walk's iteration-class totals (VG_IT_*) into
 * counts_host; synchronous. Used for algorithmic-byte accounting, not for timing. */
int vg_trace_frame_counts(vg_world *world, const vg_camera *camera_host, int32_t width,
                          int32_t height, float primary_tmax, uint32_t primary_flags,
                          const float sun[3], float shadow_tmax, vg_hit *hits,
                          uint8_t *occluded, uint64_t counts_host[VG_IT_COUNT]);
int vg_trace_rays_counts(vg_world *world, const vg_ray *rays, size_t n, uint32_t flags,
                         vg_hit *hits, uint64_t counts_host[VG_IT_COUNT]);

/* --------------------------------------------------------- dynamic world -- */
/* A sphere edit (apply_sphere, src/app.rs:199-210): every voxel at offset d from the centre
 * with |d|^2 <= radius^2 is set to `material` (0 carves) through apply_edit
 * (src/voxel.rs:450-474): skipped outside the window, stored at its toroidal slot. */
typedef struct vg_edit {
    int32_t center[3]; /* world voxel coordinates */
    uint8_t radius;    /* at most 64 */
    uint8_t material;
    uint8_t pad[2];
} vg_edit;

/* Apply n sphere edits (host array) to an uploaded world, in order: a voxel covered by
 * several spheres ends with the last one's material, as sequential set_voxel calls
 * (src/voxel.rs:871-946) leave it. The brick occupancy / materials, the brick and tile
 * uniform tables and the tile, chunk and L4 masks are updated on the device, bottom-up for
 * the changed nodes only. The window origin must be chunk-aligned (a multiple of 64), as
 * the reference's always is. Enqueued on `stream`; the edits are copied before return. */
int vg_world_apply_edits(vg_world *world, const vg_edit *edits_host, uint32_t n, void *stream);

/* Words of a frame block mask: one bit per 8x8 pixel block (row-major over
 * ceil(width/8) x ceil(height/8) blocks) plus one whole-frame flag word. */
int vg_frame_mask_words(int32_t width, int32_t height, uint32_t *words_out);
/* The reference's project_brick_to_tiles (src/temporal.rs:16-111) for the tiles the last
 * vg_world_apply_edits changed, made conservative: marks (device mask) every 8x8 block with
 * a pixel whose primary ray can enter a changed 16^3 tile, and (sun not NULL) every block
 * whose primary hit can cast a sun shadow ray (within shadow_tmax) through one. A block not
 * marked keeps its hits and shadow bytes exactly. */
int vg_frame_dirty_blocks(vg_world *world, const vg_camera *camera_host, int32_t width,
                          int32_t height, const float sun[3], float shadow_tmax, uint32_t *mask,
                          void *stream);
/* vg_trace_frame over the blocks set in a device block mask only (sun NULL: primary only);
 * pixels of other blocks are left untouched. */
int vg_trace_frame_blocks(vg_world *world, const vg_camera *camera_host, int32_t width,
                          int32_t height, float primary_tmax, uint32_t primary_flags,
                          const float sun[3], float shadow_tmax, const uint32_t *mask,
                          vg_hit *hits, uint8_t *occluded, void *stream);
/* The secondary rays shade_water_top casts at water surfaces (shaders/raymarch.wgsl:
 * 3534-3667) for a traced frame: for every pixel whose primary hit is water (materials 5..12)
 * entered through a face, with p = camera origin + d t and n the face normal, the reflection
 * ray reflect(d, n) from p + n (0.01 + 0.34 graze^2), graze = 1 - clamp(-d . n, 0, 1), and
 * the refraction ray refract(d, n, 1/1.33) (d on total internal reflection) from p + d 0.001,
 * each traced as a batched query with VG_RAY_SKIP_TRANSPARENT (trace_no_water) and tmax 200
 * (SECONDARY_MAX_T). refl_hits / refr_hits: device, one record per pixel (misses elsewhere). */
int vg_trace_water_secondary(vg_world *world, const vg_camera *camera_host, int32_t width,
                             int32_t height, const vg_hit *primary_hits, vg_hit *refl_hits,
                             vg_hit *refr_hits, void *stream);
/* The benchmark terrain (vg_generate_terrain's seeded fBm heights, caves, materials and water
 * line) generated directly in the world's device layout with its pyramid and uniform tables:
 * the same bytes as vg_generate_terrain + vg_build_pyramid + vg_world_upload at `origin`, no
 * host arrays (1.07e9 voxels of the 2048x256x2048 world never cross PCIe). Enqueued on
 * `stream`; the world counts as uploaded once this returns. */
int vg_world_generate(vg_world *world, uint64_t seed, const int32_t origin[3], void *stream);
/* The world's device state back in the reference layout (any pointer may be NULL; counts
 * from vg_world_counts). Synchronous. */
int vg_world_download(vg_world *world, vg_brick *bricks, uint64_t *tile_mask,
                      uint64_t *chunk_mask, uint64_t *l4_mask, uint8_t *brick_uniform,
                      uint8_t *tile_uniform);

/* ------------------------------------------------------------- scheduling -- */
/* Traces run as persistent, ray-compacting kernels by default: a machine-sized grid whose
 * waves refill idle lanes from a ray counter whenever fewer than `refill` of their 64 lanes
 * hold a live ray (a primary hit turns into that lane's shadow ray at once). The frame
 * kernel's 8x8 blocks are cut into VG_OPT_FRAME_STRIPES contiguous stripes, each with its
 * own counter; a wave starts on a stripe of its XCD and moves on when it is used up.
 * VG_SCHEDULE_STATIC launches one thread per ray instead. The options live in the context;
 * the autotuners time every (waves per CU, refill[, stripes]) candidate on the caller's
 * workload and keep the fastest. A world's persistent launches share one counter: launches on one world must
 * not run concurrently (use one stream per world, or synchronize). */
#define VG_SCHEDULE_PERSISTENT 0
#define VG_SCHEDULE_STATIC 1
enum {
    VG_OPT_SCHEDULE = 1,
    VG_OPT_FRAME_WAVES_PER_CU = 2,
    VG_OPT_FRAME_REFILL = 3,
    VG_OPT_RAYS_WAVES_PER_CU = 4,
    VG_OPT_RAYS_REFILL = 5,
    VG_OPT_CU_COUNT = 6, /* read only */
    VG_OPT_PROBE_WAVES_PER_CU = 7, /* probe path kernels (vg_probes_update*) */
    VG_OPT_PROBE_REFILL = 8,
    VG_OPT_PROBE_CHUNK = 9, /* probe path work ids taken per atomic: 0 (what a refill lacks) or 16..1024 */
    VG_OPT_FRAME_STRIPES = 10, /* persistent frame kernel: contiguous stripes of pixel blocks,
                                  each with its own work counter (1..64) */
    VG_OPT_FRAME_BLOCK_W = 11, /* persistent frame kernel: pixel-block width 4, 8, 16 or 32 (the
                                  block is one wave's 64 rays: height 64 / width; the beam
                                  prologue marches one box per block) */
    VG_OPT_PROBE_TBATCH = 12,  /* probe path kernel: a lane whose ray ended waits until this many
                                  of its wave's lanes have (or too few still walk) and they set
                                  up their next rays together (1..64; 1: at once) */
    VG_OPT_PROBE_MATSPEC = 13  /* probe path kernel: 1 loads each walk step's voxel material
                                  together with its pyramid words (one memory round trip per
                                  step, where a set voxel otherwise takes two; the extra load
                                  is wasted on every other step); 0 loads it when needed */
};
int vg_ctx_set_option(vg_ctx *ctx, int option, int value);
int vg_ctx_get_option(vg_ctx *ctx, int option, int *value);
int vg_autotune_frame(vg_world *world, const vg_camera *camera_host, int32_t width,
                      int32_t height, float primary_tmax, uint32_t primary_flags,
                      const float sun[3], float shadow_tmax, vg_hit *hits, uint8_t *occluded,
                      float *best_ms);
int vg_autotune_rays(vg_world *world, const vg_ray *rays, size_t n, uint32_t flags, vg_hit *hits,
                     float *best_ms);

/* ------------------------------------------------------- uneven batches -- */
/* Many batched queries at once (device descriptors): batch b traces n rays of `rays` into
 * `hits` with `flags` OR-ed into each ray. `prefix` (device, n_batches + 1 entries) holds
 * the exclusive prefix sums of the batch sizes, prefix[n_batches] == total. One persistent
 * ray-compacting launch serves all batches. */
typedef struct vg_batch {
    const vg_ray *rays;
    vg_hit *hits;
    uint32_t n;
    uint32_t flags;
} vg_batch;
int vg_trace_batches(vg_world *world, const vg_batch *batches, const uint32_t *prefix,
                     uint32_t n_batches, uint32_t total, void *stream);
int vg_trace_batches_counts(vg_world *world, const vg_batch *batches, const uint32_t *prefix,
                            uint32_t n_batches, uint32_t total,
                            uint64_t counts_host[VG_IT_COUNT]);
int vg_autotune_batches(vg_world *world, const vg_batch *batches, const uint32_t *prefix,
                        uint32_t n_batches, uint32_t total, float *best_ms);

/* ----------------------------------------------------------- HIP graphs -- */
/* Capture the launches that vg_trace_*, vg_probes_update and vg_probes_update_learned calls
 * enqueue on `stream` (not the null stream) between begin and end into an executable graph,
 * then replay it with one call. Nothing in a captured sequence allocates or memsets:
 * counters are reset by kernels, and per-call alternations (double-buffered lists) follow
 * device-side counts. A learned update with VG_NRC_TRAIN captured this way joins its
 * training step into the graph (no overlap with the next replay); join the cache
 * (vg_nrc_join) and synchronise before beginning such a capture. */
typedef struct vg_graph vg_graph;
int vg_graph_begin(vg_ctx *ctx, void *stream);
int vg_graph_end(vg_ctx *ctx, void *stream, vg_graph **out);
int vg_graph_launch(vg_graph *graph, void *stream);
int vg_graph_destroy(vg_graph *graph);

/* ------------------------------------------------------------ GI probes -- */
/* GiProbe (shaders/gi_probes.wgsl:25-48): 12 x vec4 f32 = 192 bytes. SH-L1 coefficient
 * order in each vec4 is (c0, c1y, c1z, c1x) (sh_basis, :72-74). */
typedef struct vg_probe {
    float sh_r[4], sh_g[4], sh_b[4]; /* displayed SH-L1 radiance, per channel */
    float sh_dist[4], sh_dist2[4];   /* visible distance and distance^2 moments */
    float info[4];                   /* xyz world grid coordinate held, w sample age (0 new) */
    float st_r[4], st_g[4], st_b[4]; /* staging sums of this cycle's subset estimates */
    float st_dist[4], st_dist2[4];
    float st_info[4];                /* x subset estimates this cycle, yzw relocation offset */
} vg_probe;

#define VG_GI_SKIP_BURIED 1u /* a probe whose voxel is solid traces nothing and stays empty */

/* One probe update (cs_gi_probe_update, shaders/gi_probes.wgsl:274-396). The first block
 * mirrors the shader's constants and camera fields; vg_gi_params_default derives the
 * lighting block from a sun direction with the reference's sun_color, ambient_color and
 * sky_color (shaders/common.wgsl:63-83, shaders/raymarch.wgsl:2532-2579, 2870-2876). */
typedef struct vg_gi_params {
    int32_t grid[3];        /* PROBE_DIM_X, _Y, _Z (src/world_dims.rs:44-48) */
    int32_t spacing;        /* PROBE_SPACING, voxels; probe g sits at (g + 0.5) * spacing */
    int32_t rays_per_probe; /* PROBE_RAYS: fresh gather rays per probe per update */
    int32_t epochs;         /* GI_DIR_EPOCHS: the spherical-Fibonacci set has
                               rays_per_probe * epochs points, one subset per update */
    int32_t update_div;     /* GI_UPDATE_DIV: update probes i * div + round % div */
    int32_t round;          /* camera.gi_round */
    int32_t bounces;        /* 1: ambient tail at the first hit (the reference's model);
                               2: one cosine-sampled bounce ray from the first hit */
    uint32_t seed;          /* bounce-direction sampling seed */
    uint32_t flags;         /* VG_GI_* */
    float hysteresis;       /* PROBE_FOLD_HYSTERESIS (0.6) */
    float gi_dist;          /* GI_DIST (40): reach of gather and bounce rays */
    float shadow_tmax;      /* SHADOW_MAX_DIST (480) */
    float sun[3];           /* unit vector toward the sun */
    float gi_sky;           /* GI_SKY (0.32): escaped rays gather sky * gi_sky */
    float sun_color[3];     /* sun_color(sun) */
    float ambient_tail;     /* 0.35: a last bounce adds albedo * ambient * ambient_tail */
    float ambient[3];       /* ambient_color() */
    float day_t;            /* sun_intensity(sun) */
    float sky_horizon[3];   /* mix(night, day, day_t) horizon colour */
    float dusk_w;           /* sunset warmth weight */
    float sky_zenith[3];    /* mix(night, day, day_t) zenith colour */
    float halo;             /* day_t * 0.45 */
    float sun_xz[2];        /* normalize(sun.xz) (zero when the sun is overhead) */
    float _pad[2];
} vg_gi_params;

typedef struct vg_probes vg_probes;

/* The reference's defaults (64 x 32 x 64 probes at spacing 8, 8 rays, 8 epochs, div 8,
 * one bounce) with lighting derived from `sun` (need not be normalized). */
int vg_gi_params_default(const float sun[3], vg_gi_params *out);
/* default_palette (src/renderer.rs:201-245): 256 RGBA albedo entries. */
int vg_gi_default_palette(float rgba_host[256 * 4]);
/* The gather directions of the spherical-Fibonacci set, point i = k * epochs + epoch:
 * z = 1 - 2 (i + 0.5) / n, phi = i * golden angle, d = (r cos phi, z, r sin phi) (computed
 * in double, rounded to f32): n = rays_per_probe * epochs entries of 4 floats (w = 0). */
int vg_gi_directions(int32_t n, float *xyzw_host);
/* The cosine-weighted hemisphere table bounce rays draw from: VG_GI_BOUNCE_TABLE entries
 * (t, b, n, 0) with n = cos(theta) >= 0, from a Hammersley set (computed in double). */
#define VG_GI_BOUNCE_TABLE 4096
int vg_gi_bounce_table(float *xyzw_host);

/* Probe storage for one grid (params->grid, rays_per_probe and epochs fix the sizes) on a
 * world's device. Probes start empty (age 0). */
int vg_probes_create(vg_world *world, const vg_gi_params *params, vg_probes **out);
int vg_probes_destroy(vg_probes *probes);
/* Replace the albedo palette (256 RGBA entries, host). */
int vg_probes_set_palette(vg_probes *probes, const float *rgba_host);
/* Every probe back to empty. */
int vg_probes_reset(vg_probes *probes, void *stream);
/* Device address of the PROBE_TOTAL vg_probe records (slot x + y*X + z*X*Y). */
int vg_probes_data(vg_probes *probes, vg_probe **dev_out);
/* Copy the records to / from the host. Synchronous. */
int vg_probes_download(vg_probes *probes, vg_probe *host);
int vg_probes_upload(vg_probes *probes, const vg_probe *host);
/* cs_gi_probe_update: gather rays_per_probe rays per updated probe (the round's subset),
 * shade every hit (sun with a shadow ray, plus the ambient tail or one bounce ray), project
 * onto SH-L1 with weight 4 pi / rays, fold complete cycles with hysteresis, relocate. */
int vg_probes_update(vg_probes *probes, const vg_gi_params *params, void *stream);
/* The same update reporting the traced rays (gather, shadow, bounce) and the walk's
 * iteration-class totals. Synchronous. */
int vg_probes_update_counts(vg_probes *probes, const vg_gi_params *params,
                            uint64_t rays_host[3], uint64_t counts_host[VG_IT_COUNT]);

/* ------------------------------------------------------ radiance cache -- */
/* A learned radiance cache for the probe update. It has no counterpart in the reference:
 * it stands in for the bounce tail of gather rays (the reference ends them with an ambient
 * term, shaders/gi_probes.wgsl:226-241). Position is encoded by a multiresolution hash
 * grid (per level: trilinear interpolation of 8 hashed corners, hash
 * x*1 ^ y*2654435761 ^ z*805459861 mod 2^log2_table, 2 features), concatenated with the
 * face one-hot (inputs 32..37), the direction (38..40) and a constant 1 (41) into 64
 * inputs, then three 64x64 ReLU layers and a 16x64 output layer whose rows 0..2 are the
 * rgb prediction. Inference and training run as bf16 MFMA kernels with fp32
 * accumulation; the optimizer (Adam, fp32 master parameters) skips grid entries whose
 * gradient is exactly zero. Loss: sum over queries of w * |y - t|^2 / (2 n).
 *
 * Parameter vectors (vg_nrc_get_params / set_params, fp32): mlp = W0, W1, W2 (64x64,
 * row = output, column = input) then W3 (16x64; rows 3..15 stay zero); grid =
 * [level][entry][feature], levels * 2^log2_table * 2 floats. */
#define VG_NRC_MAX_LEVELS 16
#define VG_NRC_MLP_PARAMS (3 * 64 * 64 + 16 * 64)
#define VG_NRC_MAX_BATCH (1 << 20) /* queries per vg_nrc_train_step / vg_nrc_gradients */

typedef struct vg_nrc_config {
    int32_t levels;                   /* 1..16 */
    int32_t log2_table;               /* entries per level = 2^log2_table, 4..22 */
    float scale[VG_NRC_MAX_LEVELS];   /* grid cells per voxel at each level (vg_nrc_scales) */
    float lr_grid, lr_mlp;            /* Adam learning rates */
    float beta1, beta2, eps;          /* Adam */
    float init_grid;                  /* grid features start uniform in [-init_grid, init_grid] */
    uint32_t seed;                    /* initialisation and training-sample seed */
    int32_t batch;                    /* training samples per step of the learned probe update */
    int32_t steps;                    /* training steps per learned probe update */
    int32_t _pad;
} vg_nrc_config;

/* One cache query: a surface point (world voxel coordinates), the face it lies on (low byte;
 * VG_FACE_NONE for none) and the direction that reached it. aux is carried, not read. */
typedef struct vg_nrc_query {
    float pos[3];
    uint32_t face;
    float dir[3];
    uint32_t aux;
} vg_nrc_query;

typedef struct vg_nrc vg_nrc;

/* Defaults: 16 levels of 2^19 entries, cells from 128 voxels down to 1, Adam lr 1e-2 (grid
 * and MLP), beta 0.9 / 0.99, eps 1e-10, init 1e-4, 65536 samples x 1 step per update. */
int vg_nrc_config_default(vg_nrc_config *out);
/* scale[l] = b^l / coarse_cell with b = (coarse_cell / fine_cell)^(1 / (levels - 1)). */
int vg_nrc_scales(int32_t levels, float coarse_cell, float fine_cell, float *scale_out);
/* A cache on ctx's device, parameters initialised from cfg->seed (MLP: Xavier uniform). */
int vg_nrc_create(vg_ctx *ctx, const vg_nrc_config *cfg, vg_nrc **out);
int vg_nrc_destroy(vg_nrc *nrc);
/* out[0] = MLP parameters (VG_NRC_MLP_PARAMS), out[1] = grid parameters. */
int vg_nrc_param_counts(const vg_nrc *nrc, int64_t out[2]);
/* fp32 master parameters to / from the host (either pointer may be NULL). Setting
 * parameters also resets the optimizer (moments and step count). Synchronous. */
int vg_nrc_get_params(vg_nrc *nrc, float *mlp_host, float *grid_host);
int vg_nrc_set_params(vg_nrc *nrc, const float *mlp_host, const float *grid_host);
/* Raw predictions for n device queries: out4[i] = (r, g, b, 0), device. */
int vg_nrc_infer(vg_nrc *nrc, const vg_nrc_query *queries, uint32_t n, float *out4,
                 void *stream);
/* One Adam step on n device queries with device targets (r, g, b, weight). */
int vg_nrc_train_step(vg_nrc *nrc, const vg_nrc_query *queries, const float *target4,
                      uint32_t n, void *stream);
/* Make `stream` wait for a training step still running on the cache's own stream (the
 * learned probe update enqueues its step there). The other vg_nrc_* calls and the next
 * learned update wait by themselves; a caller reading the parameters through its own
 * kernels calls this first. */
int vg_nrc_join(vg_nrc *nrc, void *stream);
/* The loss gradient of n device queries without an update (for inspection): mlp and grid
 * gradients and the loss to the host (any pointer may be NULL). Synchronous. */
int vg_nrc_gradients(vg_nrc *nrc, const vg_nrc_query *queries, const float *target4,
                     uint32_t n, float *mlp_grad_host, float *grid_grad_host, float *loss_host);
/* Optimizer steps taken and the loss of the last step. Synchronous. */
int vg_nrc_stats(vg_nrc *nrc, int64_t *steps_out, float *loss_out);
/* The last training batch the learned probe update drew (cfg batch entries): queries with
 * aux = the bounce-table index used, and targets (r, g, b, weight). Synchronous. */
int vg_nrc_last_batch(vg_nrc *nrc, vg_nrc_query *queries_host, float *target4_host);

/* The probe cache's consumer for a traced frame: indirect_light in probe mode
 * (shaders/rt_gi.wgsl:128-139) over sample_probes (shaders/gi_probes.wgsl:106-179) for every
 * pixel: at the primary hit (camera origin + dir * t, the hit face's normal, the view
 * direction), the 8 surrounding probes of the grid (params: the update's grid and spacing)
 * blended with smoothstepped trilinear weights, the backface weight and the one-tailed
 * Chebyshev visibility bound of their distance moments, each evaluated as deringed SH-L1
 * irradiance; times GI_STRENGTH (1) and the fade 1 - smoothstep(54, 90, t). Misses, hits at
 * t >= 90 and pixels with no gathered probe around get 0. hits: the frame's device records
 * (vg_trace_frame); out4: device, (r, g, b, 0) per pixel, 16-byte aligned. */
int vg_frame_indirect(vg_probes *probes, const vg_gi_params *params, const vg_camera *camera_host,
                      int32_t width, int32_t height, const vg_hit *hits, float *out4,
                      void *stream);

/* The learned probe update (the reference's cs_gi_probe_update with the cache as the
 * bounce tail): gather rays_per_probe rays per updated probe and the sun shadow ray at
 * each hit, then the tail albedo * max(0, cache(hit point, face, direction)) for every hit
 * (one inference launch over the compacted hits), the SH-L1 projection, fold and
 * relocation of vg_probes_update. With VG_NRC_TRAIN the cache then trains: cfg->steps
 * steps of cfg->batch samples drawn stratified from the previous learned update's hits,
 * each target the traced bounce tail (a cosine bounce ray and the sun shadow ray at its
 * hit; the first step's targets are traced inside this update's path launch). With one
 * step per update the step runs on the cache's own stream, overlapping the caller's next
 * work: the probe records are complete with this call's work on `stream`, the new
 * parameters after vg_nrc_join. params->bounces is not read. */
#define VG_NRC_TRAIN 1u
int vg_probes_update_learned(vg_probes *probes, const vg_gi_params *params, vg_nrc *nrc,
                             uint32_t flags, void *stream);
/* Times every schedule of the probe path kernels (waves per CU, refill threshold, work ids
 * per atomic; then the transition batch at the fastest of those) on this very update
 * (repeated; nrc may be NULL for the plain update, with VG_NRC_TRAIN the cache trains on
 * every repetition, timed as two back-to-back updates so the training step's overlap with
 * the next update counts) and keeps the fastest in the context (VG_OPT_PROBE_*). */
int vg_autotune_probes(vg_probes *probes, const vg_gi_params *params, vg_nrc *nrc,
                       uint32_t flags, float *best_ms);
/* The same, reporting rays traced (gather, shadow, training bounce), the cache queries
 * (compacted gather hits) and the path walks' iteration-class totals (may be NULL).
 * Synchronous. */
int vg_probes_update_learned_counts(vg_probes *probes, const vg_gi_params *params, vg_nrc *nrc,
                                    uint32_t flags, uint64_t rays_host[3],
                                    uint64_t *queries_host, uint64_t counts_host[VG_IT_COUNT]);
/* Per-launch timing of the probe updates that follow, with HIP events on each update's
 * stream (no counterpart in the reference; for measurement): on != 0 starts a record of up to
 * 64 updates, 0 stops and frees the events. vg_probes_timing waits for the recorded updates
 * and returns the mean milliseconds per update of [0] the prep launch, [1] the path kernel,
 * [2] the cache inference (0 in plain updates), [3] the reduction, and how many updates were
 * recorded, then starts a new record. */
int vg_probes_set_timing(vg_probes *probes, int on);
int vg_probes_timing(vg_probes *probes, float ms_out[4], int32_t *n_out);

/* raycast (src/raycast.rs:21) over a world on the device: one ray, host in and out.
 * Returns VG_OK with hit_host->hit = 0/1. Synchronous. */
int vg_raycast(vg_world *world, const float camera_world[3], const float dir[3],
               vg_hit *hit_host);

#ifdef __cplusplus
}
#endif
#endif
