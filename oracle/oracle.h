/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C restatement of voxelG's voxel ray query and the data it walks.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or as the timed CPU baseline). The
 * product path (voxelg_amd, libvoxelg_amd.so) never links, loads or calls it.
 *
 * Contents and what each follows in the reference:
 *   - world storage and index conventions    src/voxel.rs:189-326, src/world_dims.rs:9-37
 *   - pyramid / uniform-LUT builders          src/voxel.rs:233-272, 879-1000
 *   - value noise, fBm, seed offsets          src/voxel.rs:1136-1232, 1999-2021
 *   - the benchmark terrain generator         (new: fBm heightmap + 3D-noise caves)
 *   - R0: literal CPU ray query               src/raycast.rs:21-141
 *   - R1: exact hierarchical walk             shaders/raymarch.wgsl:3369-3448, 4600-4864
 *   - pinhole primary rays                    shaders/raymarch.wgsl:430-438, 600; src/camera.rs:25-42
 *   - shadow rays                             shaders/raymarch.wgsl:3953-3995, 4345-4409
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   index/bit conventions, builders  closed forms + brute-force rebuild + incremental==full
 *   noise / generator                lattice closed forms, continuity, bounds, determinism
 *   R0                               the reference's KAT scenes (src/accel.rs:515-942)
 *   R1 (LOD off)                     R1 == flat per-voxel walk; == brute-force slab test on
 *                                    tiny worlds; R1 == R0 on clean hits
 *   R1 LOD (t > 400 / 520)           hand-built bricks with known topmost material; parity
 *                                    vs the reference is UNPINNED (its shader cannot run here)
 *   primary ray generation           closed forms (centre pixel == forward, unit length)
 *   shadow                           roof-shadow KAT; any-hit == (first hit t <= tmax)
 */
#ifndef VG_ORACLE_H
#define VG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* voxelG's 72-byte Brick (src/voxel.rs:296-300): occupancy bit per voxel, voxel
 * index x + 4z + 16y, one material byte per voxel. */
typedef struct orc_brick {
    uint64_t occupancy;
    uint8_t materials[64];
} orc_brick;

/* World in the reference's own layout: x-fastest linear indices at every level. */
typedef struct orc_world {
    int32_t dims[3];      /* voxels */
    int32_t origin[3];    /* world-voxel coordinate of the window corner */
    orc_brick *bricks;    /* bx + by*BX + bz*BX*BY */
    uint64_t *tile_mask;  /* one u64 per 16^3 tile, bit per child brick */
    uint64_t *chunk_mask; /* one u64 per 64^3 chunk, bit per child tile */
    uint64_t *l4_mask;    /* one u64 per 256^3 cell, bit per child chunk */
    uint8_t *brick_uniform;
    uint8_t *tile_uniform;
    /* Every occupied voxel has a window-local y below ytop (4 x (the highest brick row that
     * holds a set tile bit + 1), a uniform tile counting as all its rows; dims[1] when the
     * window folds y): R1 ends a walk at y >= ytop that does not move down, with a miss. */
    int32_t ytop;
} orc_world;

/* Ray and hit records: the same bytes as vg_ray / vg_hit in include/vg.h. */
typedef struct orc_ray {
    float o[3];
    float tmax;
    float d[3];
    uint32_t flags;
} orc_ray;

typedef struct orc_hit {
    int32_t voxel[3];
    float t;
    uint8_t face;
    uint8_t material;
    uint8_t hit;
    uint8_t pad;
} orc_hit;

#define ORC_LOD 1u
#define ORC_SKIP_TRANSPARENT 2u
#define ORC_ANY_HIT 4u
#define ORC_BEAM 8u /* frame primary rays: fused beam prologue per 8x8 pixel block */

#define ORC_FACE_NONE 6

/* Iteration classes of the hierarchical walk, counted per ray. */
enum {
    ORC_IT_L4_SKIP = 0,
    ORC_IT_CHUNK_SKIP,
    ORC_IT_TILE_SKIP,
    ORC_IT_TILE_LOD,
    ORC_IT_TILE_UNIFORM,
    ORC_IT_BRICK_SKIP,
    ORC_IT_BRICK_UNIFORM,
    ORC_IT_BRICK_LOD,
    ORC_IT_VOXEL,
    ORC_IT_CHUNK_GROUP_SKIP, /* 2x2x2 empty chunks (128^3) */
    ORC_IT_TILE_GROUP_SKIP,  /* 2x2x2 empty tiles (32^3) */
    ORC_IT_BRICK_GROUP_SKIP, /* 2x2x2 empty bricks (8^3) */
    ORC_IT_VOXEL_GROUP_SKIP, /* 2x2x2 empty voxels (2^3) */
    ORC_IT_BEAM_TEST,        /* beam prologue: one tile tested by one lane */
    ORC_IT_COUNT
};

/* Oracle-only flag: walk the reference's levels only (no 2x2x2 group skips). */
#define ORC_NO_GROUPS 0x100u
/* Oracle-only flag: walk without the y-bound exit (orc_world.ytop). */
#define ORC_NO_YTOP 0x200u

/* dims / counts helpers */
void orc_counts(const int32_t dims[3], int64_t out[5]); /* bricks, tiles, chunks, l4, voxels */

/* storage */
int orc_world_alloc(orc_world *w, const int32_t dims[3]);
void orc_world_free(orc_world *w);
void orc_world_clear(orc_world *w);

/* builders (reference semantics) */
uint8_t orc_brick_uniform_of(const orc_brick *b);
void orc_rebuild_all_masks(orc_world *w);
void orc_rebuild_all_uniform(orc_world *w);
void orc_set_voxel(orc_world *w, int32_t x, int32_t y, int32_t z, uint8_t mat);
uint8_t orc_material_at_world(const orc_world *w, int32_t x, int32_t y, int32_t z);

/* A sphere edit (vg_edit in include/vg.h): world voxel centre, radius, material (0 carves). */
typedef struct orc_edit {
    int32_t center[3];
    uint8_t radius;
    uint8_t material;
    uint8_t pad[2];
} orc_edit;
/* apply_sphere + apply_edit (src/app.rs:199-210, src/voxel.rs:450-474) for each edit in
 * order, then a full rebuild of masks and uniform tables. */
void orc_apply_spheres(orc_world *w, const orc_edit *e, int64_t n);

/* noise + generator */
float orc_hash2(int32_t x, int32_t z);
float orc_hash3(int32_t x, int32_t y, int32_t z);
float orc_value_noise_2d(float x, float z);
float orc_value_noise_3d(float x, float y, float z);
float orc_fbm_2d(float x, float z, uint32_t octaves);
void orc_seed_offset_xz(uint64_t seed, float out[2]);
int32_t orc_terrain_height(const int32_t dims[3], uint64_t seed, int32_t x, int32_t z);
void orc_generate(orc_world *w, uint64_t seed);

/* R0: literal transliteration of src/raycast.rs, extended with tmax, t and material. */
void orc_raycast_r0(const orc_world *w, const float cam[3], const float dir[3], float tmax,
                    orc_hit *out);
/* R1: the exact hierarchical walk (counts may be NULL). */
void orc_trace_r1(const orc_world *w, const orc_ray *r, uint32_t flags, orc_hit *out,
                  uint64_t counts[ORC_IT_COUNT]);
/* R1 without any hierarchy or uniform shortcut: one voxel per iteration. */
void orc_trace_flat(const orc_world *w, const orc_ray *r, uint32_t flags, orc_hit *out);
/* Brute force over every solid voxel (tiny worlds only): nearest slab entry. */
void orc_trace_brute(const orc_world *w, const orc_ray *r, orc_hit *out, int32_t *n_ties);

/* R1 from an origin along dir exactly as given (no normalization); returns 0 when the ray
 * never enters the window (out is then a miss). */
int orc_trace_dir(const orc_world *w, const float o[3], const float dir[3], float tmax,
                  uint32_t flags, orc_hit *out, uint64_t counts[ORC_IT_COUNT]);

void orc_trace_r1_batch(const orc_world *w, const orc_ray *rays, size_t n, uint32_t flags,
                        orc_hit *out, uint64_t counts[ORC_IT_COUNT], int threads);
void orc_raycast_r0_batch(const orc_world *w, const orc_ray *rays, size_t n, orc_hit *out,
                          int threads);

/* primary rays + frame (CameraUniform-compatible 176-byte camera, see src/camera.rs:61-113) */
void orc_primary_vec(const float *camera176, int32_t px, int32_t py, float out[3]);
void orc_primary_dir(const float *camera176, int32_t px, int32_t py, float out[3]);
/* The water secondary rays of a frame's primary hits (20-byte records): two rays per pixel
 * (reflection, refraction; tmax 200, VG_RAY_SKIP_TRANSPARENT), zero directions where the
 * primary hit is not water entered through a face (shaders/raymarch.wgsl:3534-3667). */
void orc_water_rays(const float *camera176, int32_t width, int32_t height, const uint8_t *hits,
                    orc_ray *out);
void orc_primary_rays(const float *camera176, int32_t w, int32_t h, float tmax, uint32_t flags,
                      orc_ray *out);
/* One frame: primary (R1, flags) then a sun shadow ray per hit, over the rows row0,
 * row0 + row_step, ... below row1 (hits and shadow hold those rows in order), the rows
 * interleaved over the threads. */
void orc_frame(const orc_world *w, const float *camera176, int32_t width, int32_t height,
               float primary_tmax, uint32_t primary_flags, const float sun[3], float shadow_tmax,
               int32_t row0, int32_t row1, int32_t row_step, orc_hit *hits, uint8_t *shadow,
               int threads, uint64_t counts[ORC_IT_COUNT], uint32_t *iters /* may be NULL */);
/* Same frame with R0 as the traversal (the CPU baseline): raycast() is given the pinhole
 * vector and the sun direction as they are and normalizes each itself, so it traces the
 * very rays of orc_frame (and of the GPU frame). */
void orc_frame_r0(const orc_world *w, const float *camera176, int32_t width, int32_t height,
                  float primary_tmax, const float sun[3], float shadow_tmax, int32_t row0,
                  int32_t row1, int32_t row_step, orc_hit *hits, uint8_t *shadow, int threads);
void orc_shadow_origin(const float o[3], const float d[3], const orc_hit *h, float out[3]);
/* Beam prologue start t of the 8x8 pixel block (bx, by); tests counts tested tiles. */
float orc_beam_t(const orc_world *w, const float *camera176, int32_t bx, int32_t by, float tmax,
                 uint64_t *tests);

/* ---------------------------------------------------------------- GI probes --
 * FP64 restatement of the probe update (shaders/gi_probes.wgsl:187-396) over the same
 * inputs as the device: the ray origins, the traversal (R1) and every branch are computed
 * exactly as the device computes them (fp32), the radiance, SH projection and fold in
 * double. Parameters have the byte layout of vg_gi_params (include/vg.h). */
typedef struct orc_gi_params {
    int32_t grid[3];
    int32_t spacing;
    int32_t rays_per_probe;
    int32_t epochs;
    int32_t update_div;
    int32_t round;
    int32_t bounces;
    uint32_t seed;
    uint32_t flags;
    float hysteresis;
    float gi_dist;
    float shadow_tmax;
    float sun[3];
    float gi_sky;
    float sun_color[3];
    float ambient_tail;
    float ambient[3];
    float day_t;
    float sky_horizon[3];
    float dusk_w;
    float sky_zenith[3];
    float halo;
    float sun_xz[2];
    float pad[2];
} orc_gi_params;

#define ORC_GI_SKIP_BURIED 1u
#define ORC_GI_BOUNCE_TABLE 4096
#define ORC_PROBE_FIELDS 48 /* vg_probe: 12 x vec4 */

/* ------------------------------------------------------------ radiance cache --
 * The learned probe update's cache (vg_nrc_config / vg_nrc_query in include/vg.h, same
 * bytes): hash-grid encoding of the position + face one-hot + direction + a constant 1,
 * a 64-wide input, three 64x64 ReLU layers and a 16x64 output layer of which rows 0..2 are
 * the rgb prediction. Parameters are the logical fp32 vectors: mlp = W0, W1, W2 (64x64,
 * row = output) then W3 (16x64); grid[(level * 2^log2_table + entry) * 2 + feature]. */
#define ORC_NRC_WIDTH 64
#define ORC_NRC_MAX_LEVELS 16
#define ORC_NRC_IN_FACE 32 /* inputs 32..37: face one-hot */
#define ORC_NRC_IN_DIR 38  /* inputs 38..40: direction */
#define ORC_NRC_IN_ONE 41  /* input 41: constant 1 (the first layer's bias) */
#define ORC_NRC_MLP_PARAMS (3 * 64 * 64 + 16 * 64)

typedef struct orc_nrc_config {
    int32_t levels;
    int32_t log2_table;
    float scale[ORC_NRC_MAX_LEVELS];
    float lr_grid, lr_mlp;
    float beta1, beta2, eps;
    float init_grid;
    uint32_t seed;
    int32_t batch;
    int32_t steps;
    int32_t pad;
} orc_nrc_config;

typedef struct orc_nrc_query {
    float pos[3];
    uint32_t face; /* low byte: face code (6 = none) */
    float dir[3];
    uint32_t aux;
} orc_nrc_query;

typedef struct orc_nrc_eval {
    const orc_nrc_config *cfg;
    const float *mlp, *grid;
    int emulate;
} orc_nrc_eval;

double orc_bf16_round(double v);
uint32_t orc_nrc_hash(int32_t x, int32_t y, int32_t z, int32_t log2_table);
void orc_nrc_scales(int32_t levels, float coarse_cell, float fine_cell, float *out);
void orc_nrc_encode(const orc_nrc_config *c, const float *grid, const orc_nrc_query *q,
                    int emulate, double x[ORC_NRC_WIDTH]);
void orc_nrc_forward(const orc_nrc_config *c, const float *mlp, const float *grid, int64_t n,
                     const orc_nrc_query *q, int emulate, double *out3);
double orc_nrc_loss(const orc_nrc_config *c, const float *mlp, const float *grid, int64_t n,
                    const orc_nrc_query *q, const float *target4, int emulate);
double orc_nrc_loss_grad(const orc_nrc_config *c, const float *mlp, const float *grid, int64_t n,
                         const orc_nrc_query *q, const float *target4, int emulate,
                         double *g_mlp, double *g_grid);
void orc_nrc_adam(int64_t n, double *p, double *m, double *v, const double *g, double lr,
                  double beta1, double beta2, double eps, int64_t step, int sparse);

/* sky_color for a direction (radiance before GI_SKY), double. */
void orc_gi_sky(const orc_gi_params *p, const float d[3], double out[3]);
/* Bounce-table index of sample k of probe slot `probe` in `round`. */
uint32_t orc_gi_bounce_index(uint32_t seed, int32_t round, uint32_t probe, uint32_t k);
/* One gather sample from o along d: radiance rgb and visible distance (out[4]); face of
 * the gather hit (ORC_FACE_NONE if none); rays[3] += gather, shadow, bounce rays traced. */
void orc_gi_sample(const orc_world *w, const orc_gi_params *p, const float *btab,
                   const float *palette, const float o[3], const float d[3], uint32_t probe,
                   uint32_t k, double out[4], uint8_t *face, float *t_hit, uint64_t rays[3]);
/* One probe update over total = grid product records of ORC_PROBE_FIELDS doubles. */
void orc_gi_update(const orc_world *w, const orc_gi_params *p, const float *dirs,
                   const float *btab, const float *palette, double *probes, int threads,
                   uint64_t rays[3]);
/* The learned update: gather rays and the sun shadow at their hits as orc_gi_update, the
 * bounce tail of every hit taken from the cache (max(0, prediction) at the hit point
 * o + d t, its face and the gather direction) instead of a traced bounce. */
void orc_gi_update_learned(const orc_world *w, const orc_gi_params *p, const float *dirs,
                           const float *btab, const float *palette, const orc_nrc_eval *cache,
                           double *probes, int threads, uint64_t rays[3]);
/* The traced bounce tail at a cache query point (pos = the gather hit point, face): the
 * secondary origin pos + n * 0.03, the bounce direction from bounce-table entry bidx. */
void orc_gi_bounce_tail(const orc_world *w, const orc_gi_params *p, const float *btab,
                        const float *palette, const orc_nrc_query *q, uint32_t bidx,
                        double out[3], uint64_t rays[3]);
/* orc_gi_bounce_tail over n queries with their bounce-table indices; out: 3 doubles each. */
void orc_gi_bounce_tails(const orc_world *w, const orc_gi_params *p, const float *btab,
                         const float *palette, int64_t n, const orc_nrc_query *q,
                         const uint32_t *bidx, double *out, int threads);
/* The probe cache's consumer (fp32, the device's operation order): deringed SH-L1
 * irradiance of a 48-float probe record toward n (shaders/gi_probes.wgsl:86-96);
 * sample_probes at world point pw, normal n, view direction v (:106-179) over a grid of
 * records; indirect_light in probe mode (shaders/rt_gi.wgsl:128-139) for every pixel of a
 * frame's 20-byte hit records, out4: 4 floats per pixel (r, g, b, 0). */
void orc_sh_irradiance(const float *rec, const float n[3], float out[3]);
void orc_sample_probes(const float *probes, const int32_t grid[3], int32_t spacing,
                       const float pw[3], const float n[3], const float v[3], float out[3]);
void orc_frame_indirect(const float *probes, const int32_t grid[3], int32_t spacing,
                        const float *camera176, int32_t width, int32_t height,
                        const uint8_t *hits, float *out4);
/* The error reference: for each listed probe slot, an n_dirs-direction estimate (dirs: the
 * n_dirs-point Fibonacci set) of its SH-L1 radiance and distance moments from origin
 * origins[3i..3i+2], bounce samples drawn for sample k = 0..n_dirs-1 of `round`. out: 20
 * doubles per probe (r, g, b, dist, dist2 x 4 coefficients). */
void orc_gi_reference(const orc_world *w, const orc_gi_params *p, int32_t n_dirs,
                      const float *dirs, const float *btab, const float *palette,
                      const int64_t *slots, const float *origins, int64_t n, double *out,
                      int threads);

#ifdef __cplusplus
}
#endif
#endif
