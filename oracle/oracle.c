/*
 * CPU ORACLE — test infrastructure only (see oracle.h for scope and pins).
 * Built with -ffp-contract=off so every float expression rounds exactly as
 * written; no libm transcendental is used on any traced path.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ dims -- */

typedef struct dims_t {
    int32_t v[3], b[3], t[3], c[3], l4[3];
} dims_t;

static dims_t mkdims(const int32_t d[3]) {
    dims_t r;
    for (int k = 0; k < 3; ++k) {
        r.v[k] = d[k];
        r.b[k] = d[k] / 4;
        r.t[k] = r.b[k] / 4;
        r.c[k] = (r.t[k] + 3) / 4; /* src/world_dims.rs:26-29 */
        r.l4[k] = (r.c[k] + 3) / 4; /* src/world_dims.rs:34-37 */
    }
    return r;
}

void orc_counts(const int32_t dims[3], int64_t out[5]) {
    dims_t d = mkdims(dims);
    out[0] = (int64_t)d.b[0] * d.b[1] * d.b[2];
    out[1] = (int64_t)d.t[0] * d.t[1] * d.t[2];
    out[2] = (int64_t)d.c[0] * d.c[1] * d.c[2];
    out[3] = (int64_t)d.l4[0] * d.l4[1] * d.l4[2];
    out[4] = (int64_t)dims[0] * dims[1] * dims[2];
}

/* Linear index conventions: x fastest, then y, then z (src/voxel.rs:195-217, 274-277). */
static inline int64_t lin3(const int32_t n[3], int32_t x, int32_t y, int32_t z) {
    return (int64_t)x + (int64_t)y * n[0] + (int64_t)z * n[0] * n[1];
}
/* child bit inside its parent u64 at every level: lx + 4 lz + 16 ly (src/voxel.rs:190-192, 279-294) */
static inline int child_bit(int32_t lx, int32_t ly, int32_t lz) { return lx + lz * 4 + ly * 16; }

static inline int32_t pos_mod(int32_t a, int32_t b) { /* shaders/common.wgsl:49-52 */
    int32_t r = a % b;
    return r < 0 ? r + b : r;
}

/* --------------------------------------------------------------- storage -- */

int orc_world_alloc(orc_world *w, const int32_t dims[3]) {
    memset(w, 0, sizeof(*w));
    for (int k = 0; k < 3; ++k) {
        if (dims[k] <= 0 || dims[k] % 16 != 0) return -1;
        w->dims[k] = dims[k];
    }
    int64_t n[5];
    orc_counts(dims, n);
    w->bricks = (orc_brick *)calloc((size_t)n[0], sizeof(orc_brick));
    w->brick_uniform = (uint8_t *)calloc((size_t)n[0], 1);
    w->tile_mask = (uint64_t *)calloc((size_t)n[1], 8);
    w->tile_uniform = (uint8_t *)calloc((size_t)n[1], 1);
    w->chunk_mask = (uint64_t *)calloc((size_t)n[2], 8);
    w->l4_mask = (uint64_t *)calloc((size_t)n[3], 8);
    w->ytop = dims[1];
    if (!w->bricks || !w->brick_uniform || !w->tile_mask || !w->tile_uniform || !w->chunk_mask ||
        !w->l4_mask) {
        orc_world_free(w);
        return -2;
    }
    return 0;
}

void orc_world_free(orc_world *w) {
    free(w->bricks);
    free(w->brick_uniform);
    free(w->tile_mask);
    free(w->tile_uniform);
    free(w->chunk_mask);
    free(w->l4_mask);
    memset(w, 0, sizeof(*w));
}

void orc_world_clear(orc_world *w) {
    int64_t n[5];
    orc_counts(w->dims, n);
    memset(w->bricks, 0, (size_t)n[0] * sizeof(orc_brick));
    memset(w->brick_uniform, 0, (size_t)n[0]);
    memset(w->tile_mask, 0, (size_t)n[1] * 8);
    memset(w->tile_uniform, 0, (size_t)n[1]);
    memset(w->chunk_mask, 0, (size_t)n[2] * 8);
    memset(w->l4_mask, 0, (size_t)n[3] * 8);
}

/* -------------------------------------------------------------- builders -- */

/* src/voxel.rs:233-247 */
uint8_t orc_brick_uniform_of(const orc_brick *b) {
    if (b->occupancy != ~(uint64_t)0) return 0;
    uint8_t m0 = b->materials[0];
    if (m0 == 0) return 0;
    for (int i = 1; i < 64; ++i)
        if (b->materials[i] != m0) return 0;
    return m0;
}

/* src/voxel.rs:253-272 */
static uint8_t tile_uniform_of(const orc_world *w, const dims_t *d, int64_t ti) {
    int32_t tx = (int32_t)(ti % d->t[0]);
    int32_t ty = (int32_t)((ti / d->t[0]) % d->t[1]);
    int32_t tz = (int32_t)(ti / ((int64_t)d->t[0] * d->t[1]));
    int32_t bx0 = tx * 4, by0 = ty * 4, bz0 = tz * 4;
    uint8_t m0 = w->brick_uniform[lin3(d->b, bx0, by0, bz0)];
    if (m0 == 0) return 0;
    for (int dz = 0; dz < 4; ++dz)
        for (int dy = 0; dy < 4; ++dy)
            for (int dx = 0; dx < 4; ++dx)
                if (w->brick_uniform[lin3(d->b, bx0 + dx, by0 + dy, bz0 + dz)] != m0) return 0;
    return m0;
}

/* src/voxel.rs:950-960 */
void orc_rebuild_all_uniform(orc_world *w) {
    dims_t d = mkdims(w->dims);
    int64_t n[5];
    orc_counts(w->dims, n);
    for (int64_t i = 0; i < n[0]; ++i) w->brick_uniform[i] = orc_brick_uniform_of(&w->bricks[i]);
    for (int64_t i = 0; i < n[1]; ++i) w->tile_uniform[i] = tile_uniform_of(w, &d, i);
}

/* src/voxel.rs:962-1000: bottom-up OR of non-empty children. */
void orc_rebuild_all_masks(orc_world *w) {
    dims_t d = mkdims(w->dims);
    int64_t n[5];
    orc_counts(w->dims, n);
    memset(w->tile_mask, 0, (size_t)n[1] * 8);
    memset(w->chunk_mask, 0, (size_t)n[2] * 8);
    memset(w->l4_mask, 0, (size_t)n[3] * 8);
    for (int32_t bz = 0; bz < d.b[2]; ++bz)
        for (int32_t by = 0; by < d.b[1]; ++by)
            for (int32_t bx = 0; bx < d.b[0]; ++bx)
                if (w->bricks[lin3(d.b, bx, by, bz)].occupancy != 0)
                    w->tile_mask[lin3(d.t, bx / 4, by / 4, bz / 4)] |=
                        (uint64_t)1 << child_bit(bx & 3, by & 3, bz & 3);
    for (int32_t tz = 0; tz < d.t[2]; ++tz)
        for (int32_t ty = 0; ty < d.t[1]; ++ty)
            for (int32_t tx = 0; tx < d.t[0]; ++tx)
                if (w->tile_mask[lin3(d.t, tx, ty, tz)] != 0)
                    w->chunk_mask[lin3(d.c, tx / 4, ty / 4, tz / 4)] |=
                        (uint64_t)1 << child_bit(tx & 3, ty & 3, tz & 3);
    for (int32_t cz = 0; cz < d.c[2]; ++cz)
        for (int32_t cy = 0; cy < d.c[1]; ++cy)
            for (int32_t cx = 0; cx < d.c[0]; ++cx)
                if (w->chunk_mask[lin3(d.c, cx, cy, cz)] != 0)
                    w->l4_mask[lin3(d.l4, cx / 4, cy / 4, cz / 4)] |=
                        (uint64_t)1 << child_bit(cx & 3, cy & 3, cz & 3);
}

/* src/voxel.rs:879-914: propagate one brick's empty<->non-empty change upward. */
static void refresh_masks_for_brick(orc_world *w, const dims_t *d, int32_t bx, int32_t by,
                                    int32_t bz) {
    int solid = w->bricks[lin3(d->b, bx, by, bz)].occupancy != 0;
    int32_t tx = bx / 4, ty = by / 4, tz = bz / 4;
    int64_t ti = lin3(d->t, tx, ty, tz);
    uint64_t bit = (uint64_t)1 << child_bit(bx & 3, by & 3, bz & 3);
    uint64_t prev = w->tile_mask[ti];
    if (solid) w->tile_mask[ti] |= bit;
    else w->tile_mask[ti] &= ~bit;
    uint64_t now = w->tile_mask[ti];
    if ((prev == 0) != (now == 0)) {
        int32_t cx = tx / 4, cy = ty / 4, cz = tz / 4;
        int64_t ci = lin3(d->c, cx, cy, cz);
        uint64_t cbit = (uint64_t)1 << child_bit(tx & 3, ty & 3, tz & 3);
        uint64_t cprev = w->chunk_mask[ci];
        if (now == 0) w->chunk_mask[ci] &= ~cbit;
        else w->chunk_mask[ci] |= cbit;
        uint64_t cnow = w->chunk_mask[ci];
        if ((cprev == 0) != (cnow == 0)) {
            int64_t li = lin3(d->l4, cx / 4, cy / 4, cz / 4);
            uint64_t lbit = (uint64_t)1 << child_bit(cx & 3, cy & 3, cz & 3);
            if (cnow == 0) w->l4_mask[li] &= ~lbit;
            else w->l4_mask[li] |= lbit;
        }
    }
}

/* src/voxel.rs:916-935 (Brick::set :311-324); coordinates are window-local slots. */
void orc_set_voxel(orc_world *w, int32_t x, int32_t y, int32_t z, uint8_t mat) {
    if (x < 0 || y < 0 || z < 0 || x >= w->dims[0] || y >= w->dims[1] || z >= w->dims[2]) return;
    dims_t d = mkdims(w->dims);
    int32_t bx = x / 4, by = y / 4, bz = z / 4;
    int64_t bi = lin3(d.b, bx, by, bz);
    orc_brick *b = &w->bricks[bi];
    int was_empty = b->occupancy == 0;
    int vi = child_bit(x & 3, y & 3, z & 3);
    uint64_t bit = (uint64_t)1 << vi;
    if (mat == 0) {
        b->occupancy &= ~bit;
        b->materials[vi] = 0;
    } else {
        b->occupancy |= bit;
        b->materials[vi] = mat;
    }
    int is_empty = b->occupancy == 0;
    if (was_empty != is_empty) refresh_masks_for_brick(w, &d, bx, by, bz);
    w->brick_uniform[bi] = orc_brick_uniform_of(b);
    int64_t ti = lin3(d.t, bx / 4, by / 4, bz / 4);
    w->tile_uniform[ti] = tile_uniform_of(w, &d, ti);
}

/* src/voxel.rs:1136-1156: world coordinate -> toroidal slot (x, z folded). */
/* apply_sphere (src/app.rs:199-210) for each edit in order: every voxel at offset
 * (dx, dy, dz) with dx^2 + dy^2 + dz^2 <= r^2 goes through apply_edit (src/voxel.rs:450-474):
 * skipped outside the window, else stored at its toroidal slot (x, z rem_euclid the window
 * size, y relative to the window) with Brick::set semantics (src/voxel.rs:315-325). The
 * pyramid and uniform tables are then rebuilt in full, which equals the reference's
 * per-voxel refresh (test: incremental == full rebuild). */
void orc_apply_spheres(orc_world *w, const orc_edit *e, int64_t n) {
    dims_t d = mkdims(w->dims);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t r = e[i].radius, r2 = r * r;
        for (int32_t dy = -r; dy <= r; ++dy)
            for (int32_t dx = -r; dx <= r; ++dx)
                for (int32_t dz = -r; dz <= r; ++dz) {
                    if (dx * dx + dy * dy + dz * dz > r2) continue;
                    const int32_t wx = e[i].center[0] + dx, wy = e[i].center[1] + dy,
                                  wz = e[i].center[2] + dz;
                    const int32_t rx = wx - w->origin[0], ry = wy - w->origin[1],
                                  rz = wz - w->origin[2];
                    if (rx < 0 || ry < 0 || rz < 0 || rx >= w->dims[0] || ry >= w->dims[1] ||
                        rz >= w->dims[2])
                        continue;
                    const int32_t sx = pos_mod(wx, w->dims[0]), sy = ry, sz = pos_mod(wz, w->dims[2]);
                    orc_brick *b = &w->bricks[lin3(d.b, sx / 4, sy / 4, sz / 4)];
                    const int vi = child_bit(sx & 3, sy & 3, sz & 3);
                    const uint64_t bit = (uint64_t)1 << vi;
                    if (e[i].material == 0) {
                        b->occupancy &= ~bit;
                        b->materials[vi] = 0;
                    } else {
                        b->occupancy |= bit;
                        b->materials[vi] = e[i].material;
                    }
                }
    }
    orc_rebuild_all_masks(w);
    orc_rebuild_all_uniform(w);
}

uint8_t orc_material_at_world(const orc_world *w, int32_t wx, int32_t wy, int32_t wz) {
    int32_t rx = wx - w->origin[0], ry = wy - w->origin[1], rz = wz - w->origin[2];
    if (rx < 0 || ry < 0 || rz < 0 || rx >= w->dims[0] || ry >= w->dims[1] || rz >= w->dims[2])
        return 0;
    dims_t d = mkdims(w->dims);
    int32_t sx = pos_mod(wx, w->dims[0]), sy = pos_mod(wy, w->dims[1]), sz = pos_mod(wz, w->dims[2]);
    const orc_brick *b = &w->bricks[lin3(d.b, sx / 4, sy / 4, sz / 4)];
    int vi = child_bit(sx & 3, sy & 3, sz & 3);
    if (!((b->occupancy >> vi) & 1)) return 0;
    return b->materials[vi];
}

/* ----------------------------------------------------------------- noise -- */

/* src/voxel.rs:1171-1192 (u32 wrapping arithmetic) */
float orc_hash2(int32_t x, int32_t z) {
    uint32_t h = (uint32_t)x * 0x9E3779B1u + (uint32_t)z * 0x85EBCA77u;
    h = h * 0xC2B2AE3Du;
    h = h ^ (h >> 16);
    h = h * 0x85EBCA6Bu;
    return (float)(h & 0xFFFFFFu) / (float)0xFFFFFF * 2.0f - 1.0f;
}

float orc_hash3(int32_t x, int32_t y, int32_t z) {
    uint32_t h = (uint32_t)x * 0x9E3779B1u + (uint32_t)y * 0x85EBCA77u + (uint32_t)z * 0xC2B2AE3Du;
    h = h * 0xD2B74407u;
    h = h ^ (h >> 16);
    h = h * 0x85EBCA6Bu;
    return (float)(h & 0xFFFFFFu) / (float)0xFFFFFF * 2.0f - 1.0f;
}

static inline float smooth(float t) { return t * t * (3.0f - 2.0f * t); } /* src/voxel.rs:1195 */

/* src/voxel.rs:1199-1211 */
float orc_value_noise_2d(float x, float z) {
    int32_t xi = (int32_t)floorf(x), zi = (int32_t)floorf(z);
    float xf = smooth(x - (float)xi), zf = smooth(z - (float)zi);
    float v00 = orc_hash2(xi, zi), v10 = orc_hash2(xi + 1, zi);
    float v01 = orc_hash2(xi, zi + 1), v11 = orc_hash2(xi + 1, zi + 1);
    float a = v00 * (1.0f - xf) + v10 * xf;
    float b = v01 * (1.0f - xf) + v11 * xf;
    return a * (1.0f - zf) + b * zf;
}

/* src/voxel.rs:1220-1232 */
float orc_fbm_2d(float x, float z, uint32_t octaves) {
    float total = 0.0f, amp = 1.0f, freq = 1.0f, max_amp = 0.0f;
    for (uint32_t i = 0; i < octaves; ++i) {
        total += orc_value_noise_2d(x * freq, z * freq) * amp;
        max_amp += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return total / max_amp;
}

/* src/voxel.rs:1999-2021 */
float orc_value_noise_3d(float x, float y, float z) {
    int32_t xi = (int32_t)floorf(x), yi = (int32_t)floorf(y), zi = (int32_t)floorf(z);
    float xf = smooth(x - (float)xi), yf = smooth(y - (float)yi), zf = smooth(z - (float)zi);
    float v000 = orc_hash3(xi, yi, zi), v100 = orc_hash3(xi + 1, yi, zi);
    float v010 = orc_hash3(xi, yi + 1, zi), v110 = orc_hash3(xi + 1, yi + 1, zi);
    float v001 = orc_hash3(xi, yi, zi + 1), v101 = orc_hash3(xi + 1, yi, zi + 1);
    float v011 = orc_hash3(xi, yi + 1, zi + 1), v111 = orc_hash3(xi + 1, yi + 1, zi + 1);
    float a = v000 * (1.0f - xf) + v100 * xf;
    float b = v010 * (1.0f - xf) + v110 * xf;
    float c = v001 * (1.0f - xf) + v101 * xf;
    float dd = v011 * (1.0f - xf) + v111 * xf;
    float ab = a * (1.0f - yf) + b * yf;
    float cd = c * (1.0f - yf) + dd * yf;
    return ab * (1.0f - zf) + cd * zf;
}

/* src/voxel.rs:1163-1168 */
void orc_seed_offset_xz(uint64_t seed, float out[2]) {
    float hi = (float)(int32_t)(uint32_t)(seed >> 32);
    float lo = (float)(int32_t)(uint32_t)(seed & 0xFFFFFFFFu);
    out[0] = hi * 0.01734f;
    out[1] = lo * 0.02153f;
}

/* Benchmark terrain: h = 64 + 40 * fbm5((x + sx) * 0.007, (z + sz) * 0.007), clamped to
 * [2, Y-1] and truncated like `base_h.clamp(..) as i32` (src/voxel.rs:1312-1313). */
int32_t orc_terrain_height(const int32_t dims[3], uint64_t seed, int32_t x, int32_t z) {
    float s[2];
    orc_seed_offset_xz(seed, s);
    float fx = ((float)x + s[0]) * 0.007f;
    float fz = ((float)z + s[1]) * 0.007f;
    float h = 64.0f + 40.0f * orc_fbm_2d(fx, fz, 5);
    float top = (float)(dims[1] - 1);
    if (h < 2.0f) h = 2.0f;
    if (h > top) h = top;
    return (int32_t)h;
}

#define MAT_SAND 1
#define MAT_GRASS 2
#define MAT_DIRT 3
#define MAT_STONE 4
#define MAT_WATER_L8 12
#define SEA_Y 56

/* Fill the window with the benchmark terrain. World voxel (origin + local) is stored
 * at its toroidal slot. Materials by depth: surface grass (sand where h <= 57), four
 * layers of dirt, stone below; caves where value_noise_3d(x*.045, y*.085, z*.045) > 0.6
 * (not in the top 3 layers nor the bottom 5); water L8 in open air below y = 56. */
void orc_generate(orc_world *w, uint64_t seed) {
    orc_world_clear(w);
    dims_t d = mkdims(w->dims);
    for (int32_t lz = 0; lz < w->dims[2]; ++lz) {
        for (int32_t lx = 0; lx < w->dims[0]; ++lx) {
            int32_t wx = w->origin[0] + lx, wz = w->origin[2] + lz;
            int32_t sx = pos_mod(wx, w->dims[0]), sz = pos_mod(wz, w->dims[2]);
            int32_t h = orc_terrain_height(w->dims, seed, wx, wz);
            int32_t ytop = h > SEA_Y - 1 ? h : SEA_Y - 1;
            if (ytop > w->dims[1] - 1) ytop = w->dims[1] - 1;
            for (int32_t y = 0; y <= ytop; ++y) {
                uint8_t m;
                if (y > h) {
                    m = MAT_WATER_L8;
                } else {
                    if (y > 4 && y + 3 < h &&
                        orc_value_noise_3d((float)wx * 0.045f, (float)y * 0.085f,
                                           (float)wz * 0.045f) > 0.6f)
                        continue;
                    if (y == h) m = h <= 57 ? MAT_SAND : MAT_GRASS;
                    else if (y + 4 >= h) m = MAT_DIRT;
                    else m = MAT_STONE;
                }
                orc_brick *b = &w->bricks[lin3(d.b, sx / 4, y / 4, sz / 4)];
                int vi = child_bit(sx & 3, y & 3, sz & 3);
                b->occupancy |= (uint64_t)1 << vi;
                b->materials[vi] = m;
            }
        }
    }
    orc_rebuild_all_masks(w);
    orc_rebuild_all_uniform(w);
}

/* ------------------------------------------------------------ materials -- */

#define MAT_LEAF_FRINGE 33
#define MAT_TURF 35

static inline int is_transparent(uint32_t m) { /* shaders/raymarch.wgsl:891-896 */
    return (m >= 5 && m <= 12) || m == 18;
}
static inline int is_foliage(uint32_t m) { /* shaders/raymarch.wgsl:901-907 */
    return m == 14 || m == 25 || m == 26 || m == 27 || m == 30 || m == 31 || m == 33 || m == 34 ||
           m == 35 || m == 36 || m == 37;
}
static inline int uniform_optimisable(uint32_t m) { /* shaders/raymarch.wgsl:897-900 */
    return !is_foliage(m) && !is_transparent(m);
}
/* Which materials stop a ray. Every solid voxel resolves as a plain cube (the
 * src/raycast.rs and shaders/accel_probe.wgsl contract); the canopy fringe never does
 * (src/raycast.rs:88-89); skip-transparent passes water and glass (raymarch.wgsl:3510-3525);
 * any-hit (shadow) ignores fringe and turf (raymarch.wgsl:4314-4320). */
static inline int solid_mat(uint32_t m, uint32_t flags) {
    if (m == 0 || m == MAT_LEAF_FRINGE) return 0;
    if ((flags & ORC_SKIP_TRANSPARENT) && is_transparent(m)) return 0;
    if ((flags & ORC_ANY_HIT) && m == MAT_TURF) return 0;
    return 1;
}

static inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }

/* shaders/raymarch.wgsl:253-295: topmost non-turf solid voxel's material, y layer 3 first. */
static uint32_t brick_topmost_material(const orc_brick *b) {
    for (int y = 3; y >= 0; --y) {
        uint64_t layer = (b->occupancy >> (16 * y)) & 0xFFFFu;
        if (layer) {
            uint32_t m = b->materials[16 * y + ctz64(layer)];
            if (m != MAT_TURF) return m;
        }
    }
    return b->occupancy ? MAT_TURF : 0;
}

/* ------------------------------------------------------------- R0 (raycast.rs) -- */

static inline float safe_inv(float x) { return fabsf(x) < 1e-8f ? 1e30f : 1.0f / x; }
static inline float fmin_(float a, float b) { return a < b ? a : b; }
static inline float fmax_(float a, float b) { return a > b ? a : b; }

static void miss(orc_hit *o) {
    memset(o, 0, sizeof(*o));
    o->t = -1.0f;
    o->face = ORC_FACE_NONE;
}

/* glam 0.29 Vec3::normalize_or_zero: self * (1 / sqrt(x*x + y*y + z*z)), zero unless the
 * reciprocal is finite and positive. Returns 0 for a zero result. */
static int normalize_or_zero(const float in[3], float out[3]) {
    float len = sqrtf(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
    float rcp = 1.0f / len;
    if (!(isfinite(rcp) && rcp > 0.0f)) return 0;
    out[0] = in[0] * rcp;
    out[1] = in[1] * rcp;
    out[2] = in[2] * rcp;
    return !(out[0] == 0.0f && out[1] == 0.0f && out[2] == 0.0f);
}

static int r0_in_window(const orc_world *w, const int32_t v[3]) {
    for (int k = 0; k < 3; ++k) {
        int32_t r = v[k] - w->origin[k];
        if (r < 0 || r >= w->dims[k]) return 0;
    }
    return 1;
}

static int r0_in_solid(const orc_world *w, const dims_t *d, const int32_t v[3]) {
    if (!r0_in_window(w, v)) return 0;
    /* rem_euclid on x and z, y as is (src/raycast.rs:74-76); with the window's y origin at
     * 0, the only one the reference uses, pos_mod on y is the identity */
    int32_t sx = pos_mod(v[0], w->dims[0]), sy = pos_mod(v[1], w->dims[1]);
    int32_t sz = pos_mod(v[2], w->dims[2]);
    const orc_brick *b = &w->bricks[lin3(d->b, sx / 4, sy / 4, sz / 4)];
    if (b->occupancy == 0) return 0;
    int vi = child_bit(sx & 3, sy & 3, sz & 3);
    if (!((b->occupancy >> vi) & 1)) return 0;
    return b->materials[vi] != MAT_LEAF_FRINGE;
}

/* src/accel.rs:578-586 */
static float ray_box_entry(const float o[3], const float dd[3], const int32_t v[3]) {
    float tm = -INFINITY;
    for (int k = 0; k < 3; ++k) {
        float inv = safe_inv(dd[k]);
        float a = ((float)v[k] - o[k]) * inv;
        float b = ((float)v[k] + 1.0f - o[k]) * inv;
        tm = fmax_(tm, fmin_(a, b));
    }
    return fmax_(tm, 0.0f);
}

/* src/raycast.rs:21-133 line by line; extensions: entry-t of the current voxel is carried
 * so a per-ray tmax can end the walk, the hit reports ray_box_entry t and the material. */
void orc_raycast_r0(const orc_world *w, const float cam[3], const float dir_in[3], float tmax,
                    orc_hit *out) {
    miss(out);
    float dir[3];
    if (!normalize_or_zero(dir_in, dir)) return;
    dims_t dm = mkdims(w->dims);
    float win_min[3], win_max[3], inv[3], t0[3], t1[3];
    for (int k = 0; k < 3; ++k) {
        win_min[k] = (float)w->origin[k];
        win_max[k] = win_min[k] + (float)w->dims[k];
        inv[k] = safe_inv(dir[k]);
        t0[k] = (win_min[k] - cam[k]) * inv[k];
        t1[k] = (win_max[k] - cam[k]) * inv[k];
    }
    float tmn[3], tmx[3];
    for (int k = 0; k < 3; ++k) {
        tmn[k] = fmin_(t0[k], t1[k]);
        tmx[k] = fmax_(t0[k], t1[k]);
    }
    float t_enter = fmax_(fmax_(tmn[0], fmax_(tmn[1], tmn[2])), 0.0f);
    float t_exit = fmin_(tmx[0], fmin_(tmx[1], tmx[2]));
    if (t_enter >= t_exit || t_exit < 0.0f) return;

    int32_t voxel[3], step[3];
    float t_delta[3], t_max[3];
    for (int k = 0; k < 3; ++k) {
        float p = cam[k] + dir[k] * (t_enter + 1e-3f);
        float lo = win_min[k] + 1e-3f, hi = win_max[k] - 1e-3f;
        p = fmin_(fmax_(p, lo), hi);
        voxel[k] = (int32_t)floorf(p);
        step[k] = dir[k] > 0.0f ? 1 : (dir[k] < 0.0f ? -1 : 0);
        t_delta[k] = fabsf(inv[k]);
    }
    for (int k = 0; k < 3; ++k) {
        float nxt = step[k] > 0 ? (float)(voxel[k] + 1) : (float)voxel[k];
        t_max[k] = (nxt - cam[k]) * inv[k];
    }
    int last_axis = -1, steps_left = 4096;
    float t_cur = t_enter;

#define R0_STEP(AX_OUT)                                                        \
    do {                                                                       \
        int ax_ = -1;                                                          \
        float m_ = INFINITY;                                                   \
        for (int k_ = 0; k_ < 3; ++k_)                                         \
            if (step[k_] != 0 && t_max[k_] < m_) {                             \
                ax_ = k_;                                                      \
                m_ = t_max[k_];                                                \
            }                                                                  \
        if (ax_ >= 0) {                                                        \
            voxel[ax_] += step[ax_];                                           \
            t_cur = t_max[ax_];                                                \
            t_max[ax_] += t_delta[ax_];                                        \
        }                                                                      \
        AX_OUT = ax_;                                                          \
    } while (0)

    /* Phase 1: leave any solid voxels the ray starts inside (src/raycast.rs:108-115). */
    while (steps_left > 0 && r0_in_solid(w, &dm, voxel)) {
        if (t_cur > tmax) return;
        int ax;
        R0_STEP(ax);
        if (ax < 0) return;
        last_axis = ax;
        steps_left--;
        if (!r0_in_window(w, voxel)) return;
    }
    /* Phase 2: first solid voxel (src/raycast.rs:117-131). */
    while (steps_left > 0) {
        if (!r0_in_window(w, voxel)) return;
        if (t_cur > tmax) return;
        if (r0_in_solid(w, &dm, voxel)) {
            out->hit = 1;
            out->voxel[0] = voxel[0];
            out->voxel[1] = voxel[1];
            out->voxel[2] = voxel[2];
            out->face = last_axis >= 0 ? (uint8_t)(2 * last_axis + (step[last_axis] > 0 ? 1 : 0))
                                       : ORC_FACE_NONE;
            out->t = ray_box_entry(cam, dir, voxel);
            out->material = orc_material_at_world(w, voxel[0], voxel[1], voxel[2]);
            return;
        }
        int ax;
        R0_STEP(ax);
        if (ax < 0) return;
        last_axis = ax;
        steps_left--;
    }
#undef R0_STEP
}

/* ---------------------------------------------------- R1 (exact hierarchical) -- */

typedef struct r1_t {
    const orc_world *w;
    dims_t d;
    float org[3], dir[3], inv[3], tmin3[3];
    int32_t step[3], l[3], ro_mod[3];
    int32_t ytop;
    float t_cur;
    int last_axis;
} r1_t;

/* Direct-form t of the plane a ray crosses to ENTER voxel e along axis k. */
static inline float t_enter_voxel(const r1_t *s, int k, int32_t e) {
    int32_t plane = s->step[k] > 0 ? e : e + 1;
    return ((float)plane - s->org[k]) * s->inv[k];
}

/* Ray setup (shaders/raymarch.wgsl:3369-3411 with src/raycast.rs's 1e-3 clamp).
 * dir must be the already-normalized direction. Returns 0 for a window miss. */
static int r1_init(r1_t *s, const orc_world *w, const float o[3], const float dir[3]) {
    s->w = w;
    s->d = mkdims(w->dims);
    float t0[3], t1[3], tmx[3];
    for (int k = 0; k < 3; ++k) {
        s->org[k] = o[k] - (float)w->origin[k];
        s->dir[k] = dir[k];
        s->inv[k] = safe_inv(dir[k]);
        t0[k] = (0.0f - s->org[k]) * s->inv[k];
        t1[k] = ((float)w->dims[k] - s->org[k]) * s->inv[k];
        s->tmin3[k] = fmin_(t0[k], t1[k]);
        tmx[k] = fmax_(t0[k], t1[k]);
        s->ro_mod[k] = pos_mod(w->origin[k], w->dims[k]);
        /* an axis whose |d| < 1e-8 never steps (its safe_inv is the 1e30 sentinel) */
        s->step[k] = fabsf(dir[k]) < 1e-8f ? 0 : (dir[k] > 0.0f ? 1 : -1);
    }
    if (s->step[0] == 0 && s->step[1] == 0 && s->step[2] == 0) return 0; /* cannot advance */
    float t_enter = fmax_(fmax_(s->tmin3[0], fmax_(s->tmin3[1], s->tmin3[2])), 0.0f);
    float t_exit = fmin_(tmx[0], fmin_(tmx[1], tmx[2]));
    if (t_enter >= t_exit || t_exit < 0.0f) return 0;
    for (int k = 0; k < 3; ++k) {
        float p = s->org[k] + dir[k] * (t_enter + 1e-3f);
        float hi = (float)w->dims[k] - 1e-3f;
        p = fmin_(fmax_(p, 1e-3f), hi);
        s->l[k] = (int32_t)floorf(p);
    }
    s->t_cur = t_enter;
    s->last_axis = -1;
    s->ytop = w->ytop;
    return 1;
}

/* One voxel advance: the axis whose next plane is nearest, ties to the lowest axis
 * (src/raycast.rs:92-103); t recomputed from the integer plane, never accumulated. */
static void r1_step(r1_t *s) {
    int ax = -1;
    float m = 0.0f;
    for (int k = 0; k < 3; ++k) {
        if (s->step[k] == 0) continue;
        float t = t_enter_voxel(s, k, s->l[k] + s->step[k]);
        if (ax < 0 || t < m) {
            ax = k;
            m = t;
        }
    }
    s->l[ax] += s->step[ax];
    s->t_cur = m;
    s->last_axis = ax;
}

/* Exact empty-cell skip (the role of shaders/raymarch.wgsl:4807-4864). The per-voxel walk
 * processes plane crossings in (t, axis) order; it leaves the aligned cell at the first
 * cell-face crossing (T*, ax*). Every other axis has then crossed exactly the in-cell planes
 * ordered before (T*, ax*); that count is found from a floor estimate and corrected with
 * the same direct-form t comparisons, so the result equals the per-voxel walk's state. */
static void r1_skip(r1_t *s, const int32_t slot[3], int32_t S) {
    int32_t lo[3];
    int ax = -1;
    float ts = 0.0f;
    for (int k = 0; k < 3; ++k) {
        lo[k] = s->l[k] - (slot[k] & (S - 1));
        if (s->step[k] == 0) continue;
        int32_t plane = s->step[k] > 0 ? lo[k] + S : lo[k];
        float t = ((float)plane - s->org[k]) * s->inv[k];
        if (ax < 0 || t < ts) {
            ts = t;
            ax = k;
        }
    }
    int32_t nl[3];
    for (int k = 0; k < 3; ++k) {
        if (k == ax) {
            nl[k] = s->step[k] > 0 ? lo[k] + S : lo[k] - 1;
            continue;
        }
        if (s->step[k] == 0) {
            nl[k] = s->l[k];
            continue;
        }
        int32_t st = s->step[k], first = s->l[k], last = st > 0 ? lo[k] + S - 1 : lo[k];
        int32_t e = (int32_t)floorf(s->org[k] + s->dir[k] * ts);
        if (st > 0) {
            if (e < first) e = first;
            if (e > last) e = last;
        } else {
            if (e > first) e = first;
            if (e < last) e = last;
        }
#define BEFORE(T) ((T) < ts || ((T) == ts && k < ax))
        while (e != first && !BEFORE(t_enter_voxel(s, k, e))) e -= st;
        while (e != last && BEFORE(t_enter_voxel(s, k, e + st))) e += st;
#undef BEFORE
        nl[k] = e;
    }
    s->l[0] = nl[0];
    s->l[1] = nl[1];
    s->l[2] = nl[2];
    s->t_cur = ts;
    s->last_axis = ax;
}

/* Jump the walk to the state reached once every plane crossing with t < tb has been
 * processed (a prefix of the (t, axis) order): each axis lands on its last plane before tb
 * (floor estimate corrected by direct t comparisons), the entry t is the latest crossing,
 * ties to the higher axis (the one processed last). Used after the beam prologue proved
 * every voxel before tb empty, so the walk continues exactly as if it had stepped there. */
static void r1_advance_to(r1_t *s, float tb) {
    float tl = -INFINITY;
    int la = -1;
    for (int k = 0; k < 3; ++k) {
        int32_t st = s->step[k];
        if (st == 0 || !(t_enter_voxel(s, k, s->l[k] + st) < tb)) continue;
        int32_t nearest = s->l[k] + st;
        int32_t e = (int32_t)floorf(s->org[k] + s->dir[k] * tb);
        if (st > 0 ? e < nearest : e > nearest) e = nearest;
        while (e != nearest && !(t_enter_voxel(s, k, e) < tb)) e -= st;
        while (t_enter_voxel(s, k, e + st) < tb) e += st;
        float tk = t_enter_voxel(s, k, e);
        s->l[k] = e;
        if (tk > tl || (tk == tl && k > la)) {
            tl = tk;
            la = k;
        }
    }
    if (la >= 0) {
        s->t_cur = tl;
        s->last_axis = la;
    }
}

static void r1_emit(const r1_t *s, uint32_t mat, orc_hit *out) {
    out->hit = 1;
    for (int k = 0; k < 3; ++k) out->voxel[k] = s->l[k] + s->w->origin[k];
    out->t = s->t_cur;
    out->material = (uint8_t)mat;
    int la = s->last_axis;
    if (la < 0) { /* began inside the cell: dominant entry plane (raymarch.wgsl:3336-3340) */
        la = (s->tmin3[0] >= s->tmin3[1] && s->tmin3[0] >= s->tmin3[2]) ? 0
             : (s->tmin3[1] >= s->tmin3[2])                               ? 1
                                                                          : 2;
    }
    out->face = s->step[la] > 0 ? (uint8_t)(2 * la + 1)
                : s->step[la] < 0 ? (uint8_t)(2 * la)
                                  : ORC_FACE_NONE;
    out->pad = 0;
}

/* The 8 child bits of the 2x2x2 group of children holding slot coordinate sl inside a
 * parent u64 whose children are 2^(sh-1) voxels wide: bits lx + 4 lz + 16 ly for lx, ly, lz
 * in the group's octant. An empty group is an aligned empty cube twice the child size,
 * skipped exactly like any other empty cell. */
static inline uint64_t group_mask(const int32_t sl[3], int sh) {
    int ox = (sl[0] >> sh) & 1, oy = (sl[1] >> sh) & 1, oz = (sl[2] >> sh) & 1;
    return (uint64_t)0x330033u << (2 * ox + 8 * oz + 32 * oy);
}

#define LOD_BRICK_T 400.0f /* shaders/raymarch.wgsl:4489 */
#define TILE_LOD_T 520.0f  /* shaders/raymarch.wgsl:4494 */

/* The walk itself. Phase 1 (raycast.rs) leaves solid voxels the ray starts in; phase 2
 * returns the first solid voxel. Each iteration tests the coarsest level first and skips
 * the largest empty aligned cell (raymarch.wgsl:4626-4775). Uniform cells end phase 2 at
 * once (the entry voxel is solid) and are skipped whole in phase 1. LOD shortcuts
 * (raymarch.wgsl:4673-4758) apply in phase 2 of plain first-hit queries only. */
static void r1_walk(r1_t *s, float tmax, uint32_t flags, orc_hit *out, uint64_t *cnt) {
    const orc_world *w = s->w;
    const dims_t *d = &s->d;
    const int lod = (flags & ORC_LOD) && !(flags & (ORC_SKIP_TRANSPARENT | ORC_ANY_HIT));
    const int groups = !(flags & ORC_NO_GROUPS);
    const int32_t ytop = (flags & ORC_NO_YTOP) ? w->dims[1] : s->ytop;
    int phase1 = 1;
    for (;;) {
        if (s->t_cur > tmax) return;
        if (s->l[0] < 0 || s->l[1] < 0 || s->l[2] < 0 || s->l[0] >= w->dims[0] ||
            s->l[1] >= w->dims[1] || s->l[2] >= w->dims[2])
            return;
        /* above every occupied voxel and not moving down: nothing left to hit */
        if (s->step[1] >= 0 && s->l[1] >= ytop) return;
        int32_t sl[3];
        for (int k = 0; k < 3; ++k) {
            sl[k] = s->l[k] + s->ro_mod[k];
            if (sl[k] >= w->dims[k]) sl[k] -= w->dims[k];
        }
        uint64_t l4w = w->l4_mask[lin3(d->l4, sl[0] >> 8, sl[1] >> 8, sl[2] >> 8)];
        if (l4w == 0) {
            if (cnt) cnt[ORC_IT_L4_SKIP]++;
            phase1 = 0;
            r1_skip(s, sl, 256);
            continue;
        }
        if (!((l4w >> child_bit((sl[0] >> 6) & 3, (sl[1] >> 6) & 3, (sl[2] >> 6) & 3)) & 1)) {
            phase1 = 0;
            if (groups && !(l4w & group_mask(sl, 7))) {
                if (cnt) cnt[ORC_IT_CHUNK_GROUP_SKIP]++;
                r1_skip(s, sl, 128);
            } else {
                if (cnt) cnt[ORC_IT_CHUNK_SKIP]++;
                r1_skip(s, sl, 64);
            }
            continue;
        }
        uint64_t cw = w->chunk_mask[lin3(d->c, sl[0] >> 6, sl[1] >> 6, sl[2] >> 6)];
        if (!((cw >> child_bit((sl[0] >> 4) & 3, (sl[1] >> 4) & 3, (sl[2] >> 4) & 3)) & 1)) {
            phase1 = 0;
            if (groups && !(cw & group_mask(sl, 5))) {
                if (cnt) cnt[ORC_IT_TILE_GROUP_SKIP]++;
                r1_skip(s, sl, 32);
            } else {
                if (cnt) cnt[ORC_IT_TILE_SKIP]++;
                r1_skip(s, sl, 16);
            }
            continue;
        }
        int64_t ti = lin3(d->t, sl[0] >> 4, sl[1] >> 4, sl[2] >> 4);
        uint64_t tw = w->tile_mask[ti];
        if (lod && !phase1 && s->t_cur > TILE_LOD_T) {
            /* tile_representative_material: first occupied child brick's top (:4782-4798) */
            if (cnt) cnt[ORC_IT_TILE_LOD]++;
            if (tw) {
                int lin = ctz64(tw);
                int32_t bx = (sl[0] >> 4) * 4 + (lin & 3);
                int32_t bz = (sl[2] >> 4) * 4 + ((lin >> 2) & 3);
                int32_t by = (sl[1] >> 4) * 4 + ((lin >> 4) & 3);
                uint32_t lm = brick_topmost_material(&w->bricks[lin3(d->b, bx, by, bz)]);
                if (lm != 0 && uniform_optimisable(lm)) {
                    r1_emit(s, lm, out);
                    return;
                }
            }
        }
        uint32_t tu = w->tile_uniform[ti];
        if (tu != 0) {
            if (cnt) cnt[ORC_IT_TILE_UNIFORM]++;
            if (solid_mat(tu, flags)) {
                if (!phase1) {
                    r1_emit(s, tu, out);
                    return;
                }
            } else {
                phase1 = 0;
            }
            r1_skip(s, sl, 16);
            continue;
        }
        if (!((tw >> child_bit((sl[0] >> 2) & 3, (sl[1] >> 2) & 3, (sl[2] >> 2) & 3)) & 1)) {
            phase1 = 0;
            /* inside an occupied tile a tile-LOD test may fire at any visited voxel past
             * t 520, so a group skip there must end before t 520: its exit lies within the
             * 8^3 cell's diagonal (< 14) of the current voxel's entry */
            if (groups && !(tw & group_mask(sl, 3)) &&
                (!lod || s->t_cur + 14.0f <= TILE_LOD_T)) {
                if (cnt) cnt[ORC_IT_BRICK_GROUP_SKIP]++;
                r1_skip(s, sl, 8);
            } else {
                if (cnt) cnt[ORC_IT_BRICK_SKIP]++;
                r1_skip(s, sl, 4);
            }
            continue;
        }
        int64_t bi = lin3(d->b, sl[0] >> 2, sl[1] >> 2, sl[2] >> 2);
        uint32_t bu = w->brick_uniform[bi];
        if (bu != 0) {
            if (cnt) cnt[ORC_IT_BRICK_UNIFORM]++;
            if (solid_mat(bu, flags)) {
                if (!phase1) {
                    r1_emit(s, bu, out);
                    return;
                }
            } else {
                phase1 = 0;
            }
            r1_skip(s, sl, 4);
            continue;
        }
        const orc_brick *b = &w->bricks[bi];
        if (lod && !phase1 && s->t_cur > LOD_BRICK_T) {
            if (cnt) cnt[ORC_IT_BRICK_LOD]++;
            uint32_t m = brick_topmost_material(b);
            if (!is_foliage(m)) {
                r1_emit(s, m, out);
                return;
            }
        }
        int vi = child_bit(sl[0] & 3, sl[1] & 3, sl[2] & 3);
        if ((b->occupancy >> vi) & 1) {
            if (cnt) cnt[ORC_IT_VOXEL]++;
            uint32_t m = b->materials[vi];
            if (solid_mat(m, flags)) {
                if (!phase1) {
                    r1_emit(s, m, out);
                    return;
                }
                r1_step(s);
                continue;
            }
            phase1 = 0;
            r1_step(s);
            continue;
        }
        phase1 = 0;
        /* an empty 2x2x2 voxel octant is left whole; inside an occupied brick a brick-LOD
         * test may fire at any visited voxel past t 400, so the octant (diagonal < 4) must
         * be left before it */
        if (groups && !(b->occupancy & group_mask(sl, 1)) &&
            (!lod || s->t_cur + 4.0f <= LOD_BRICK_T)) {
            if (cnt) cnt[ORC_IT_VOXEL_GROUP_SKIP]++;
            r1_skip(s, sl, 2);
        } else {
            if (cnt) cnt[ORC_IT_VOXEL]++;
            r1_step(s);
        }
    }
}

static void trace_core_from(const orc_world *w, const float o[3], const float dir[3], float tmax,
                            uint32_t flags, float t_start, orc_hit *out, uint64_t *cnt) {
    miss(out);
    r1_t s;
    if (!r1_init(&s, w, o, dir)) return;
    if (t_start > 0.0f && s.t_cur <= tmax) r1_advance_to(&s, t_start);
    r1_walk(&s, tmax, flags, out, cnt);
}

static void trace_core(const orc_world *w, const float o[3], const float dir[3], float tmax,
                       uint32_t flags, orc_hit *out, uint64_t *cnt) {
    trace_core_from(w, o, dir, tmax, flags, 0.0f, out, cnt);
}

int orc_trace_dir(const orc_world *w, const float o[3], const float dir[3], float tmax,
                  uint32_t flags, orc_hit *out, uint64_t counts[ORC_IT_COUNT]) {
    miss(out);
    r1_t s;
    if (!r1_init(&s, w, o, dir)) return 0;
    r1_walk(&s, tmax, flags, out, counts);
    return 1;
}

void orc_trace_r1(const orc_world *w, const orc_ray *r, uint32_t flags, orc_hit *out,
                  uint64_t counts[ORC_IT_COUNT]) {
    float dir[3];
    miss(out);
    if (!normalize_or_zero(r->d, dir)) return;
    trace_core(w, r->o, dir, r->tmax, flags | r->flags, out, counts);
}

void orc_trace_flat(const orc_world *w, const orc_ray *r, uint32_t flags, orc_hit *out) {
    float dir[3];
    miss(out);
    flags |= r->flags;
    if (!normalize_or_zero(r->d, dir)) return;
    r1_t s;
    if (!r1_init(&s, w, r->o, dir)) return;
    int phase1 = 1;
    for (;;) {
        if (s.t_cur > r->tmax) return;
        if (s.l[0] < 0 || s.l[1] < 0 || s.l[2] < 0 || s.l[0] >= w->dims[0] ||
            s.l[1] >= w->dims[1] || s.l[2] >= w->dims[2])
            return;
        int32_t wv[3] = {s.l[0] + w->origin[0], s.l[1] + w->origin[1], s.l[2] + w->origin[2]};
        uint32_t m = orc_material_at_world(w, wv[0], wv[1], wv[2]);
        if (solid_mat(m, flags)) {
            if (!phase1) {
                r1_emit(&s, m, out);
                return;
            }
        } else {
            phase1 = 0;
        }
        r1_step(&s);
    }
}

/* Brute force: nearest slab entry over every solid voxel of the window. n_ties counts solid
 * voxels whose entry lies within 1e-3 of the nearest (edge / corner ambiguity). */
void orc_trace_brute(const orc_world *w, const orc_ray *r, orc_hit *out, int32_t *n_ties) {
    miss(out);
    *n_ties = 0;
    float dir[3];
    if (!normalize_or_zero(r->d, dir)) return;
    float inv[3];
    for (int k = 0; k < 3; ++k) inv[k] = 1.0f / dir[k];
    double best = INFINITY;
    int32_t bv[3] = {0, 0, 0};
    for (int32_t z = 0; z < w->dims[2]; ++z)
        for (int32_t y = 0; y < w->dims[1]; ++y)
            for (int32_t x = 0; x < w->dims[0]; ++x) {
                int32_t v[3] = {x + w->origin[0], y + w->origin[1], z + w->origin[2]};
                uint32_t m = orc_material_at_world(w, v[0], v[1], v[2]);
                if (!solid_mat(m, r->flags)) continue;
                double tn = -INFINITY, tf = INFINITY;
                for (int k = 0; k < 3; ++k) {
                    if (dir[k] == 0.0f) {
                        if (r->o[k] < (float)v[k] || r->o[k] >= (float)v[k] + 1.0f) tn = INFINITY;
                        continue;
                    }
                    double a = ((double)v[k] - r->o[k]) * (1.0 / (double)dir[k]);
                    double b = ((double)v[k] + 1.0 - r->o[k]) * (1.0 / (double)dir[k]);
                    double lo = a < b ? a : b, hi = a < b ? b : a;
                    if (lo > tn) tn = lo;
                    if (hi < tf) tf = hi;
                }
                if (tn > tf || tf < 0.0 || tn > r->tmax) continue;
                if (tn < 0.0) tn = 0.0;
                if (tn < best - 1e-3) {
                    best = tn;
                    bv[0] = v[0];
                    bv[1] = v[1];
                    bv[2] = v[2];
                    *n_ties = 1;
                } else if (tn < best + 1e-3) {
                    (*n_ties)++;
                    if (tn < best) {
                        best = tn;
                        bv[0] = v[0];
                        bv[1] = v[1];
                        bv[2] = v[2];
                    }
                }
            }
    (void)inv;
    if (best == INFINITY) return;
    out->hit = 1;
    out->voxel[0] = bv[0];
    out->voxel[1] = bv[1];
    out->voxel[2] = bv[2];
    out->t = (float)best;
    out->material = orc_material_at_world(w, bv[0], bv[1], bv[2]);
}

/* ------------------------------------------------------------- threading -- */

typedef struct job_t {
    int kind; /* 0 = R1 batch, 1 = R0 batch, 2 = frame R1, 3 = frame R0 */
    int tid, nthreads;
    const orc_world *w;
    const orc_ray *rays;
    size_t n;
    uint32_t flags;
    orc_hit *out;
    uint64_t counts[ORC_IT_COUNT];
    /* frame */
    const float *cam;
    int32_t width, height, row0, row1, row_step; /* rows row0, row0 + step, ... < row1 */
    float ptmax, stmax, sun[3];
    float sun_in[3]; /* R0 frame: the sun direction as given (raycast normalizes it) */
    uint8_t *shadow;
    uint32_t *iters; /* optional: per pixel (primary, shadow) iteration counts */
} job_t;

static void frame_rows(job_t *j);

static void *job_run(void *arg) {
    job_t *j = (job_t *)arg;
    if (j->kind == 0) {
        for (size_t i = (size_t)j->tid; i < j->n; i += (size_t)j->nthreads)
            orc_trace_r1(j->w, &j->rays[i], j->flags, &j->out[i], j->counts);
    } else if (j->kind == 1) {
        for (size_t i = (size_t)j->tid; i < j->n; i += (size_t)j->nthreads)
            orc_raycast_r0(j->w, j->rays[i].o, j->rays[i].d, j->rays[i].tmax, &j->out[i]);
    } else {
        frame_rows(j);
    }
    return NULL;
}

static void run_jobs(job_t *proto, int threads, uint64_t *counts) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    job_t *jobs = (job_t *)calloc((size_t)threads, sizeof(job_t));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        jobs[t] = *proto;
        jobs[t].tid = t;
        jobs[t].nthreads = threads;
        memset(jobs[t].counts, 0, sizeof(jobs[t].counts));
    }
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, job_run, &jobs[t]);
    job_run(&jobs[0]);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    if (counts)
        for (int t = 0; t < threads; ++t)
            for (int k = 0; k < ORC_IT_COUNT; ++k) counts[k] += jobs[t].counts[k];
    free(jobs);
    free(th);
}

void orc_trace_r1_batch(const orc_world *w, const orc_ray *rays, size_t n, uint32_t flags,
                        orc_hit *out, uint64_t counts[ORC_IT_COUNT], int threads) {
    job_t p;
    memset(&p, 0, sizeof(p));
    p.kind = 0;
    p.w = w;
    p.rays = rays;
    p.n = n;
    p.flags = flags;
    p.out = out;
    run_jobs(&p, threads, counts);
}

void orc_raycast_r0_batch(const orc_world *w, const orc_ray *rays, size_t n, orc_hit *out,
                          int threads) {
    job_t p;
    memset(&p, 0, sizeof(p));
    p.kind = 1;
    p.w = w;
    p.rays = rays;
    p.n = n;
    p.out = out;
    run_jobs(&p, threads, NULL);
}

/* ---------------------------------------------------------------- frames -- */

/* Camera floats (src/camera.rs:61-113): origin 0..2, forward 4..6, right 8..10, up 12..14,
 * tan_half_fov 15, resolution 16..17, world_origin (i32) 20..22, jitter 24..25. */
/* The pinhole vector before normalization: forward + right * ndc.x * tan(fov/2) * aspect +
 * up * ndc.y * tan(fov/2), ndc = (2u - 1, 1 - 2v), uv = (pixel + 0.5 + jitter) / resolution
 * (raymarch.wgsl:600; ray_dir_uv :430-438). */
void orc_primary_vec(const float *cam, int32_t px, int32_t py, float r[3]) {
    float u = ((float)px + 0.5f + cam[24]) / cam[16];
    float v = ((float)py + 0.5f + cam[25]) / cam[17];
    float nx = u * 2.0f - 1.0f, ny = 1.0f - v * 2.0f;
    float aspect = cam[16] / cam[17];
    float a = nx * cam[15] * aspect, b = ny * cam[15];
    for (int k = 0; k < 3; ++k) r[k] = cam[4 + k] + cam[8 + k] * a + cam[12 + k] * b;
}

/* The primary ray direction: the pinhole vector normalized as glam's normalize_or_zero does
 * (v * (1 / sqrt(v . v))), so raycast() given the pinhole vector traces this very direction. */
void orc_primary_dir(const float *cam, int32_t px, int32_t py, float out[3]) {
    float r[3];
    orc_primary_vec(cam, px, py, r);
    float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    float rcp = 1.0f / len;
    out[0] = r[0] * rcp;
    out[1] = r[1] * rcp;
    out[2] = r[2] * rcp;
}

void orc_primary_rays(const float *cam, int32_t width, int32_t height, float tmax, uint32_t flags,
                      orc_ray *out) {
    for (int32_t y = 0; y < height; ++y)
        for (int32_t x = 0; x < width; ++x) {
            orc_ray *r = &out[(size_t)y * width + x];
            r->o[0] = cam[0];
            r->o[1] = cam[1];
            r->o[2] = cam[2];
            r->tmax = tmax;
            orc_primary_dir(cam, x, y, r->d);
            r->flags = flags;
        }
}

/* Shadow origin: p_hit + n * 0.001 with n the hit face normal (raymarch.wgsl:3955). */
void orc_shadow_origin(const float o[3], const float d[3], const orc_hit *h, float out[3]) {
    for (int k = 0; k < 3; ++k) {
        float n = 0.0f;
        if (h->face < 6 && h->face / 2 == k) n = (h->face & 1) ? -1.0f : 1.0f;
        float p = o[k] + d[k] * h->t;
        out[k] = p + n * 0.001f;
    }
}

#define BEAM_STEP 16.0f
#define BEAM_MAX_STEPS 64

/* Beam prologue of one 8x8 pixel block (the role of cs_beam, shaders/beam.wgsl:31-113, made
 * conservative): the 64 rays share the camera origin O, so on t in [t0, t1] they all lie in
 * the box O + [min d, max d] * [t0, t1] (per axis, over the block's directions). March that
 * box in 16-voxel steps, grown by one voxel, and test every world tile it touches inside the
 * window (a 4x4x4 cube at most, one tile per lane of a wave) against the chunk masks; the block's start
 * t is the first step whose box touches an occupied tile. Every voxel a block ray visits
 * before it is empty, so the walk may jump there exactly (r1_advance_to). */
float orc_beam_t(const orc_world *w, const float *cam, int32_t bx, int32_t by, float tmax,
                 uint64_t *tests) {
    dims_t d = mkdims(w->dims);
    for (int k = 0; k < 3; ++k) {
        float rel = cam[k] - (float)w->origin[k];
        if (!(rel >= 0.0f && rel < (float)w->dims[k])) return 0.0f; /* camera outside */
    }
    float dmin[3] = {INFINITY, INFINITY, INFINITY}, dmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int32_t ly = 0; ly < 8; ++ly)
        for (int32_t lx = 0; lx < 8; ++lx) {
            float dd[3];
            orc_primary_dir(cam, bx * 8 + lx, by * 8 + ly, dd);
            for (int k = 0; k < 3; ++k) {
                dmin[k] = fmin_(dmin[k], dd[k]);
                dmax[k] = fmax_(dmax[k], dd[k]);
            }
        }
    /* every block ray rises (or stays level) and a step's box lies at or above the world's
     * occupied y bound: nothing is left to hit, as when the box leaves the window */
    const int rising = dmin[1] >= 0.0f;
    const float ybound = (float)(w->origin[1] + w->ytop);
    float t0 = 0.0f;
    for (int st = 0; st < BEAM_MAX_STEPS && t0 < tmax; ++st) {
        float t1 = (float)(st + 1) * BEAM_STEP;
        int32_t lo[3], hi[3];
        int64_t n = 1;
        float mny = 0.0f;
        for (int k = 0; k < 3; ++k) {
            float a = cam[k] + dmin[k] * t0, b = cam[k] + dmin[k] * t1;
            float c = cam[k] + dmax[k] * t0, e = cam[k] + dmax[k] * t1;
            float mn = fmin_(fmin_(a, b), fmin_(c, e)) - 1.0f;
            float mx = fmax_(fmax_(a, b), fmax_(c, e)) + 1.0f;
            if (k == 1) mny = mn;
            int32_t tl = (int32_t)floorf(mn * 0.0625f), th = (int32_t)floorf(mx * 0.0625f);
            /* world tiles of the window */
            int32_t wl = w->origin[k] >> 4, wh = (w->origin[k] + w->dims[k] - 1) >> 4;
            if (tl < wl) tl = wl;
            if (th > wh) th = wh;
            lo[k] = tl;
            hi[k] = th;
            n *= th >= tl ? (int64_t)(th - tl + 1) : 0;
        }
        if (n == 0 || (rising && mny >= ybound)) {
            /* the box left the window (every block ray has exited it for good), or every
             * block ray is above every occupied voxel from here on */
            t0 = t1;
            break;
        }
        if (hi[0] - lo[0] >= 4 || hi[1] - lo[1] >= 4 || hi[2] - lo[2] >= 4) break; /* > 4x4x4 */
        int occupied = 0;
        for (int64_t j = 0; j < n; ++j) {
            int32_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
            int32_t tw[3] = {lo[0] + (int32_t)(j % nx), lo[1] + (int32_t)((j / nx) % ny),
                             lo[2] + (int32_t)(j / ((int64_t)nx * ny))};
            int32_t sl[3];
            for (int k = 0; k < 3; ++k) sl[k] = pos_mod(tw[k] * 16, w->dims[k]);
            uint64_t cw = w->chunk_mask[lin3(d.c, sl[0] >> 6, sl[1] >> 6, sl[2] >> 6)];
            if ((cw >> child_bit((sl[0] >> 4) & 3, (sl[1] >> 4) & 3, (sl[2] >> 4) & 3)) & 1)
                occupied = 1;
        }
        if (tests) *tests += (uint64_t)n;
        if (occupied) break;
        t0 = t1;
    }
    return t0;
}

static void frame_rows(job_t *j) {
    const float *cam = j->cam;
    const int32_t step = j->row_step > 0 ? j->row_step : 1;
    const int32_t nrows = j->row1 > j->row0 ? (j->row1 - j->row0 + step - 1) / step : 0;
    /* the selected rows interleaved over the threads */
    for (int32_t k = j->tid; k < nrows; k += j->nthreads) {
        const int32_t y = j->row0 + k * step;
        for (int32_t x = 0; x < j->width; ++x) {
            size_t i = (size_t)k * j->width + x;
            float dir[3];
            orc_primary_dir(cam, x, y, dir);
            orc_hit *h = &j->out[i];
            uint64_t c[ORC_IT_COUNT] = {0};
            float tb = 0.0f;
            if (j->kind == 2 && (j->flags & ORC_BEAM)) {
                /* the block's beam start; its tile tests are charged to the block's first pixel */
                uint64_t *tc = ((x & 7) == 0 && (y & 7) == 0) ? &c[ORC_IT_BEAM_TEST] : NULL;
                tb = orc_beam_t(j->w, cam, x >> 3, y >> 3, j->ptmax, tc);
            }
            if (j->kind == 2) {
                trace_core_from(j->w, cam, dir, j->ptmax, j->flags, tb, h, c);
            } else {
                /* raycast() normalizes its direction itself: given the pinhole vector it
                 * traces exactly the primary direction above */
                float pv[3];
                orc_primary_vec(cam, x, y, pv);
                orc_raycast_r0(j->w, cam, pv, j->ptmax, h);
            }
            uint32_t it_primary = 0;
            for (int k = 0; k < ORC_IT_COUNT; ++k) it_primary += (uint32_t)c[k];
            uint8_t occ = 0;
            if (h->hit) {
                float so[3];
                orc_shadow_origin(cam, dir, h, so);
                orc_hit sh;
                if (j->kind == 2)
                    trace_core(j->w, so, j->sun, j->stmax, ORC_ANY_HIT | (j->flags & (ORC_NO_GROUPS | ORC_NO_YTOP)),
                               &sh, c);
                else orc_raycast_r0(j->w, so, j->sun_in, j->stmax, &sh); /* normalizes it */
                occ = sh.hit;
            }
            j->shadow[i] = occ;
            uint32_t it_all = 0;
            for (int k = 0; k < ORC_IT_COUNT; ++k) {
                it_all += (uint32_t)c[k];
                j->counts[k] += c[k];
            }
            if (j->iters) {
                j->iters[2 * i] = it_primary;
                j->iters[2 * i + 1] = it_all - it_primary;
            }
        }
    }
}

void orc_frame(const orc_world *w, const float *cam, int32_t width, int32_t height,
               float primary_tmax, uint32_t primary_flags, const float sun[3], float shadow_tmax,
               int32_t row0, int32_t row1, int32_t row_step, orc_hit *hits, uint8_t *shadow,
               int threads, uint64_t counts[ORC_IT_COUNT], uint32_t *iters) {
    job_t p;
    memset(&p, 0, sizeof(p));
    p.kind = 2;
    p.w = w;
    p.cam = cam;
    p.width = width;
    p.height = height;
    p.row0 = row0;
    p.row1 = row1;
    p.row_step = row_step;
    p.ptmax = primary_tmax;
    p.stmax = shadow_tmax;
    p.flags = primary_flags;
    if (!normalize_or_zero(sun, p.sun)) p.sun[0] = p.sun[1] = p.sun[2] = 0.0f;
    p.out = hits;
    p.shadow = shadow;
    p.iters = iters;
    run_jobs(&p, threads, counts);
}

void orc_frame_r0(const orc_world *w, const float *cam, int32_t width, int32_t height,
                  float primary_tmax, const float sun[3], float shadow_tmax, int32_t row0,
                  int32_t row1, int32_t row_step, orc_hit *hits, uint8_t *shadow, int threads) {
    job_t p;
    memset(&p, 0, sizeof(p));
    p.kind = 3;
    p.w = w;
    p.cam = cam;
    p.width = width;
    p.height = height;
    p.row0 = row0;
    p.row1 = row1;
    p.row_step = row_step;
    p.ptmax = primary_tmax;
    p.stmax = shadow_tmax;
    for (int k = 0; k < 3; ++k) p.sun_in[k] = sun[k];
    if (!normalize_or_zero(sun, p.sun)) p.sun[0] = p.sun[1] = p.sun[2] = 0.0f;
    p.out = hits;
    p.shadow = shadow;
    run_jobs(&p, threads, NULL);
}

/* The secondary rays shade_water_top casts at a water surface (shaders/raymarch.wgsl:
 * 3534-3667), fp32 in the device's operation order: for a primary hit of water (5..12)
 * entered through a face, p = origin + d t, n the face normal; reflection reflect(d, n)
 * (WGSL: d - 2 dot(n, d) n) from p + n (0.01 + 0.34 graze^2), graze = 1 - clamp(-d . n, 0, 1);
 * refraction refract(d, n, 1/1.33) (WGSL: k = 1 - eta^2 (1 - dot^2), zero if k < 0, else
 * eta d - (eta dot + sqrt k) n; d itself when its length is below 0.01) from p + d 0.001.
 * out: two rays per pixel (reflection, refraction), tmax 200, VG_RAY_SKIP_TRANSPARENT; a
 * pixel without a water hit gets zero-direction rays (they miss). */
void orc_water_rays(const float *cam, int32_t width, int32_t height, const uint8_t *hits,
                    orc_ray *out) {
    for (int64_t pix = 0; pix < (int64_t)width * height; ++pix) {
        const uint8_t *hr = hits + 20 * pix;
        float t;
        memcpy(&t, hr + 12, 4);
        const uint32_t face = hr[16], mat = hr[17], hit = hr[18];
        orc_ray *a = out + 2 * pix, *b = a + 1;
        memset(a, 0, 2 * sizeof(orc_ray));
        a->tmax = b->tmax = 200.0f;
        a->flags = b->flags = 2u;
        if (!(hit && mat >= 5u && mat <= 12u && face < 6u)) continue;
        float d[3], n[3], pw[3];
        orc_primary_dir(cam, (int32_t)(pix % width), (int32_t)(pix / width), d);
        for (int k = 0; k < 3; ++k) {
            n[k] = ((int)(face >> 1) == k) ? ((face & 1u) ? -1.0f : 1.0f) : 0.0f;
            pw[k] = cam[k] + d[k] * t;
        }
        const float dn = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        float c = -d[0] * n[0] + -d[1] * n[1] + -d[2] * n[2];
        c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
        const float graze = 1.0f - c;
        const float lift = 0.01f + 0.34f * graze * graze;
        for (int k = 0; k < 3; ++k) {
            a->d[k] = d[k] - 2.0f * dn * n[k];
            a->o[k] = pw[k] + n[k] * lift;
        }
        const float eta = (float)(1.0 / 1.33);
        const float kk = 1.0f - eta * eta * (1.0f - dn * dn);
        float f[3] = {0.0f, 0.0f, 0.0f};
        if (!(kk < 0.0f)) {
            const float sk = sqrtf(kk);
            for (int k = 0; k < 3; ++k) f[k] = eta * d[k] - (eta * dn + sk) * n[k];
        }
        const int tir = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) < 0.01f;
        for (int k = 0; k < 3; ++k) {
            b->d[k] = tir ? d[k] : f[k];
            b->o[k] = pw[k] + d[k] * 0.001f;
        }
    }
}
