// Host-side world building for the engine: error plumbing, window geometry, the
// benchmark terrain generator, the pyramid / uniform-LUT builders and the camera
// basis. Everything here runs on the CPU before upload; the traversal itself only
// runs on the GPU (vg_device.hip).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <thread>
#include <vector>

#include "vg_internal.h"

namespace vg {

static thread_local std::string g_err;

int fail(int status, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return status;
}

int make_dims(const int32_t dims[3], Dims *o) {
    if (!dims || !o) return fail(VG_E_INVALID, "null dims");
    for (int k = 0; k < 3; ++k) {
        if (dims[k] <= 0 || dims[k] % 64 != 0 || dims[k] > 16384)
            return fail(VG_E_INVALID, "dims[%d]=%d must be a positive multiple of 64 <= 16384", k,
                        dims[k]);
        o->v[k] = dims[k];
        o->b[k] = dims[k] / 4;
        o->t[k] = dims[k] / 16;
        o->c[k] = dims[k] / 64;
        o->l4[k] = (o->c[k] + 3) / 4;
    }
    o->n_bricks = (int64_t)o->b[0] * o->b[1] * o->b[2];
    o->n_tiles = (int64_t)o->t[0] * o->t[1] * o->t[2];
    o->n_chunks = (int64_t)o->c[0] * o->c[1] * o->c[2];
    o->n_l4 = (int64_t)o->l4[0] * o->l4[1] * o->l4[2];
    return VG_OK;
}

int resolve_threads(int threads) {
    if (threads > 0) return std::min(threads, 256);
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? (int)std::min(hc, 256u) : 1;
}

// Run fn(i) for i in [0, n) over `threads` workers with dynamic chunking.
template <class F> static void parallel_for(int64_t n, int threads, int64_t grain, F fn) {
    threads = (int)std::min<int64_t>(threads, std::max<int64_t>(1, (n + grain - 1) / grain));
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            int64_t i0 = next.fetch_add(grain);
            if (i0 >= n) break;
            int64_t i1 = std::min(n, i0 + grain);
            for (int64_t i = i0; i < i1; ++i) fn(i);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
}

// ------------------------------------------------------------------ noise --
// Value noise and fBm as in src/voxel.rs:1171-1232 and :1999-2021: u32 wrapping
// hashes, smoothstep-weighted lerps, fp32 throughout (built with contraction off).

static inline float hash2(int32_t x, int32_t z) {
    uint32_t h = (uint32_t)x * 0x9E3779B1u + (uint32_t)z * 0x85EBCA77u;
    h *= 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    return (float)(h & 0xFFFFFFu) / (float)0xFFFFFF * 2.0f - 1.0f;
}

static inline float hash3(int32_t x, int32_t y, int32_t z) {
    uint32_t h = (uint32_t)x * 0x9E3779B1u + (uint32_t)y * 0x85EBCA77u + (uint32_t)z * 0xC2B2AE3Du;
    h *= 0xD2B74407u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    return (float)(h & 0xFFFFFFu) / (float)0xFFFFFF * 2.0f - 1.0f;
}

static inline float smoothw(float t) { return t * t * (3.0f - 2.0f * t); }

static inline float lerpw(float a, float b, float w) { return a * (1.0f - w) + b * w; }

static float noise2(float x, float z) {
    int32_t xi = (int32_t)std::floor(x), zi = (int32_t)std::floor(z);
    float xf = smoothw(x - (float)xi), zf = smoothw(z - (float)zi);
    float a = lerpw(hash2(xi, zi), hash2(xi + 1, zi), xf);
    float b = lerpw(hash2(xi, zi + 1), hash2(xi + 1, zi + 1), xf);
    return lerpw(a, b, zf);
}

static float noise3(float x, float y, float z) {
    int32_t xi = (int32_t)std::floor(x), yi = (int32_t)std::floor(y), zi = (int32_t)std::floor(z);
    float xf = smoothw(x - (float)xi), yf = smoothw(y - (float)yi), zf = smoothw(z - (float)zi);
    float a = lerpw(hash3(xi, yi, zi), hash3(xi + 1, yi, zi), xf);
    float b = lerpw(hash3(xi, yi + 1, zi), hash3(xi + 1, yi + 1, zi), xf);
    float c = lerpw(hash3(xi, yi, zi + 1), hash3(xi + 1, yi, zi + 1), xf);
    float d = lerpw(hash3(xi, yi + 1, zi + 1), hash3(xi + 1, yi + 1, zi + 1), xf);
    return lerpw(lerpw(a, b, yf), lerpw(c, d, yf), zf);
}

static float fbm2(float x, float z, int octaves) {
    float total = 0.0f, amp = 1.0f, freq = 1.0f, max_amp = 0.0f;
    for (int i = 0; i < octaves; ++i) {
        total += noise2(x * freq, z * freq) * amp;
        max_amp += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return total / max_amp;
}

struct TerrainParams {
    float sx, sz, top;
};

static TerrainParams terrain_params(const int32_t dims[3], uint64_t seed) {
    // seed_offset_xz, src/voxel.rs:1163-1168
    TerrainParams p;
    p.sx = (float)(int32_t)(uint32_t)(seed >> 32) * 0.01734f;
    p.sz = (float)(int32_t)(uint32_t)(seed & 0xFFFFFFFFu) * 0.02153f;
    p.top = (float)(dims[1] - 1);
    return p;
}

// Surface height: 64 + 40 * fBm5 at 0.007 voxel^-1, clamped to [2, Y-1], truncated.
static int32_t height_at(const TerrainParams &p, int32_t x, int32_t z) {
    float h = 64.0f + 40.0f * fbm2(((float)x + p.sx) * 0.007f, ((float)z + p.sz) * 0.007f, 5);
    h = std::min(std::max(h, 2.0f), p.top);
    return (int32_t)h;
}

} // namespace vg

using namespace vg;

extern "C" {

int vg_abi_version(void) { return VG_ABI_VERSION; }

const char *vg_last_error(void) { return g_err.c_str(); }

const char *vg_status_string(int s) {
    switch (s) {
    case VG_OK: return "ok";
    case VG_E_INVALID: return "invalid argument";
    case VG_E_NOMEM: return "out of memory";
    case VG_E_DEVICE: return "device error";
    case VG_E_UNSUPPORTED: return "unsupported";
    case VG_E_STATE: return "bad state";
    default: return "unknown status";
    }
}

int vg_world_counts(const int32_t dims[3], int64_t out[5]) {
    Dims d;
    int st = make_dims(dims, &d);
    if (st) return st;
    if (!out) return fail(VG_E_INVALID, "null out");
    out[0] = d.n_bricks;
    out[1] = d.n_tiles;
    out[2] = d.n_chunks;
    out[3] = d.n_l4;
    out[4] = (int64_t)dims[0] * dims[1] * dims[2];
    return VG_OK;
}

int vg_terrain_height(const int32_t dims[3], uint64_t seed, int32_t x, int32_t z,
                      int32_t *h_out) {
    Dims d;
    int st = make_dims(dims, &d);
    if (st) return st;
    if (!h_out) return fail(VG_E_INVALID, "null h_out");
    *h_out = height_at(terrain_params(dims, seed), x, z);
    return VG_OK;
}

int vg_generate_terrain(const int32_t dims[3], const int32_t origin[3], uint64_t seed,
                        vg_brick *bricks, int threads) {
    Dims d;
    int st = make_dims(dims, &d);
    if (st) return st;
    if (!bricks || !origin) return fail(VG_E_INVALID, "null bricks/origin");
    const TerrainParams tp = terrain_params(dims, seed);
    const int32_t Y = dims[1];
    std::memset(bricks, 0, (size_t)d.n_bricks * sizeof(vg_brick));
    // One work item per slot brick-row along z: its 4 voxel rows write disjoint bricks.
    parallel_for(d.b[2], resolve_threads(threads), 1, [&](int64_t sbz) {
        for (int32_t sz = (int32_t)sbz * 4; sz < (int32_t)sbz * 4 + 4; ++sz) {
            // world z whose slot is sz
            int32_t wz = origin[2] + pos_mod(sz - pos_mod(origin[2], dims[2]), dims[2]);
            for (int32_t sx = 0; sx < dims[0]; ++sx) {
                int32_t wx = origin[0] + pos_mod(sx - pos_mod(origin[0], dims[0]), dims[0]);
                int32_t h = height_at(tp, wx, wz);
                int32_t ytop = std::min(std::max(h, 55), Y - 1);
                for (int32_t y = 0; y <= ytop; ++y) {
                    uint8_t m;
                    if (y > h) {
                        m = 12; // water L8 in open air below y = 56
                    } else {
                        if (y > 4 && y + 3 < h &&
                            noise3((float)wx * 0.045f, (float)y * 0.085f, (float)wz * 0.045f) >
                                0.6f)
                            continue; // cave
                        if (y == h) m = h <= 57 ? 1 : 2; // sand / grass
                        else if (y + 4 >= h) m = 3;      // dirt
                        else m = 4;                      // stone
                    }
                    vg_brick &b = bricks[lin3(d.b, sx >> 2, y >> 2, sz >> 2)];
                    int vi = child_bit(sx & 3, y & 3, sz & 3);
                    b.occupancy |= (uint64_t)1 << vi;
                    b.materials[vi] = m;
                }
            }
        }
    });
    return VG_OK;
}

// brick_uniform_of (src/voxel.rs:233-247)
static inline uint8_t brick_uniform(const vg_brick &b) {
    if (b.occupancy != ~(uint64_t)0) return 0;
    uint8_t m0 = b.materials[0];
    if (!m0) return 0;
    for (int i = 1; i < 64; ++i)
        if (b.materials[i] != m0) return 0;
    return m0;
}

int vg_build_pyramid(const int32_t dims[3], const vg_brick *bricks, uint64_t *tile_mask,
                     uint64_t *chunk_mask, uint64_t *l4_mask, uint8_t *bu, uint8_t *tu,
                     int threads) {
    Dims d;
    int st = make_dims(dims, &d);
    if (st) return st;
    if (!bricks || !tile_mask || !chunk_mask || !l4_mask || !bu || !tu)
        return fail(VG_E_INVALID, "null array");
    const int nt = resolve_threads(threads);
    parallel_for(d.n_bricks, nt, 4096, [&](int64_t i) { bu[i] = brick_uniform(bricks[i]); });
    // tiles: mask = OR over non-empty child bricks; uniform iff all 64 children share one
    // uniform material (src/voxel.rs:253-272, 962-979)
    parallel_for(d.n_tiles, nt, 256, [&](int64_t ti) {
        int32_t tx = (int32_t)(ti % d.t[0]);
        int32_t ty = (int32_t)((ti / d.t[0]) % d.t[1]);
        int32_t tz = (int32_t)(ti / ((int64_t)d.t[0] * d.t[1]));
        uint64_t m = 0;
        uint8_t u0 = bu[lin3(d.b, tx * 4, ty * 4, tz * 4)];
        bool uni = u0 != 0;
        for (int dz = 0; dz < 4; ++dz)
            for (int dy = 0; dy < 4; ++dy)
                for (int dx = 0; dx < 4; ++dx) {
                    int64_t bi = lin3(d.b, tx * 4 + dx, ty * 4 + dy, tz * 4 + dz);
                    if (bricks[bi].occupancy) m |= (uint64_t)1 << child_bit(dx, dy, dz);
                    if (bu[bi] != u0) uni = false;
                }
        tile_mask[ti] = m;
        tu[ti] = uni ? u0 : 0;
    });
    parallel_for(d.n_chunks, nt, 16, [&](int64_t ci) {
        int32_t cx = (int32_t)(ci % d.c[0]);
        int32_t cy = (int32_t)((ci / d.c[0]) % d.c[1]);
        int32_t cz = (int32_t)(ci / ((int64_t)d.c[0] * d.c[1]));
        uint64_t m = 0;
        for (int dz = 0; dz < 4; ++dz)
            for (int dy = 0; dy < 4; ++dy)
                for (int dx = 0; dx < 4; ++dx)
                    if (tile_mask[lin3(d.t, cx * 4 + dx, cy * 4 + dy, cz * 4 + dz)])
                        m |= (uint64_t)1 << child_bit(dx, dy, dz);
        chunk_mask[ci] = m;
    });
    std::memset(l4_mask, 0, (size_t)d.n_l4 * 8);
    for (int32_t cz = 0; cz < d.c[2]; ++cz)
        for (int32_t cy = 0; cy < d.c[1]; ++cy)
            for (int32_t cx = 0; cx < d.c[0]; ++cx)
                if (chunk_mask[lin3(d.c, cx, cy, cz)])
                    l4_mask[lin3(d.l4, cx / 4, cy / 4, cz / 4)] |= (uint64_t)1
                                                                   << child_bit(cx & 3, cy & 3, cz & 3);
    return VG_OK;
}

// glam Vec3::normalize: v * (1 / sqrt(dot(v, v)))
static void normalize3(float v[3]) {
    float r = 1.0f / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= r;
    v[1] *= r;
    v[2] *= r;
}

int vg_camera_from_pose(const float pos[3], float yaw, float pitch, float fov_y, int32_t width,
                        int32_t height, const int32_t world_origin[3], vg_camera *out) {
    if (!pos || !out || width <= 0 || height <= 0)
        return fail(VG_E_INVALID, "bad camera arguments");
    std::memset(out, 0, sizeof(*out));
    float f[3] = {std::sin(yaw) * std::cos(pitch), std::sin(pitch), std::cos(yaw) * std::cos(pitch)};
    normalize3(f);
    float r[3] = {f[2], 0.0f, -f[0]};
    normalize3(r);
    float u[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
    normalize3(u);
    for (int k = 0; k < 3; ++k) {
        out->origin[k] = pos[k];
        out->forward[k] = f[k];
        out->right[k] = r[k];
        out->up[k] = u[k];
        out->prev_origin[k] = pos[k];
        out->prev_forward[k] = f[k];
        out->prev_right[k] = r[k];
        out->prev_up[k] = u[k];
        out->world_origin[k] = world_origin ? world_origin[k] : 0;
    }
    out->tan_half_fov = std::tan(fov_y * 0.5f);
    out->resolution[0] = (float)width;
    out->resolution[1] = (float)height;
    return VG_OK;
}

} // extern "C"
