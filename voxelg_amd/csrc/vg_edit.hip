// Dynamic world on gfx950: sphere edits of an uploaded world, applied on the device with the
// reference's sequential semantics (apply_sphere -> apply_edit -> set_voxel, src/app.rs:199-210,
// src/voxel.rs:450-474, 871-946), an incremental bottom-up rebuild of the pyramid and uniform
// tables, and the 8x8 screen blocks a frame must retrace afterwards (the reference's
// project_brick_to_tiles, src/temporal.rs:16-111, made conservative).
//
//   pairs   one wave per edit enumerates the bricks its sphere touches inside the window
//           (lane per brick of the clipped box): a count pass (atomic per-brick counters),
//           an inclusive scan of the counters, and a fill pass that places each (brick, edit)
//           pair in its brick's run, counting the counter back down to zero. The order inside
//           a run is arbitrary: the brick pass takes, per voxel, the covering edit of the
//           HIGHEST index, which is what sequential calls leave.
//           The count pass also lists the touched tiles (first pair of a brick -> first
//           brick of a tile -> one append), so later passes visit only those.
//   bricks  one 1024-thread block per touched tile, one wave per touched brick, one lane per
//           voxel: the lanes load up to 64 of the brick's edits at once and every lane tests
//           them all (operands broadcast with readlane); the occupancy word is the wave's
//           ballot, the brick-uniform byte a wave-wide equality; changed bricks mark their
//           tile dirty.
//   tiles / chunks / L4   one wave per dirty node (tiles: from the touched-tile list): the node word is the ballot of its 64
//           children being non-empty, the tile-uniform byte a wave-wide equality.
//   blocks  one wave per changed tile projects the tile's box (primary rays) and the box
//           swept against the sun over the shadow range (shadow rays) onto the screen and
//           marks the covered 8x8 blocks (conservative: 2-pixel margins, the whole screen
//           when a box straddles the camera plane).
#include "vg_world.h"

#include "vg_scan.h"

#include <algorithm>
#include <vector>

struct vg_edit_state {
    size_t cap_edits, cap_pairs;
    vg_edit *edits;          // device copy of the last batch
    uint32_t *cnt;           // per nested brick: pairs of the batch (zero between batches)
    uint32_t *end;           // inclusive scan of cnt: a brick's run is [end[bn-1], end[bn])
    uint32_t *list;          // edit index of every (brick, edit) pair, grouped by brick
    void *temp;
    size_t temp_bytes;
    uint32_t *tile_flag;     // per nested tile: listed in this batch (zero between batches)
    uint32_t *tile_list;     // nested tiles holding a brick some edit of the batch touches
    uint32_t *tile_n;        // [0] entries of tile_list
    uint32_t *tile_dirty;    // bit per nested tile changed by the last batch
    uint32_t *chunk_dirty;   // bit per chunk
    uint32_t *l4_dirty;      // bit per L4 cell
    uint32_t *blist;         // dirty screen blocks of a masked frame
    uint32_t *blist_n;
    size_t blist_cap;
    vg_edit *staging[2];     // pinned host staging of the edit batch, double-buffered
    hipEvent_t staged[2];    // the copy out of staging[i] has finished
    bool staged_rec[2];
    int slot;
};

namespace vg {

constexpr int kMaxEditRadius = 64;

// Slot coordinate -> the window's world coordinate on one axis.
__device__ __forceinline__ int32_t slot_to_world(int32_t s, int32_t origin, int32_t dim) {
    int32_t r = (s - origin) % dim;
    r = r < 0 ? r + dim : r;
    return origin + r;
}

// Nested brick index -> slot brick coordinates.
__device__ __forceinline__ void brick_slot(const DevWorld &w, uint32_t bn, int32_t b[3]) {
    const uint32_t tn = bn >> 6, ci = tn >> 6;
    const uint32_t bb = bn & 63u, tb = tn & 63u;
    const int32_t cx = (int32_t)(ci % (uint32_t)w.cdim[0]);
    const int32_t cy = (int32_t)((ci / (uint32_t)w.cdim[0]) % (uint32_t)w.cdim[1]);
    const int32_t cz = (int32_t)(ci / ((uint32_t)w.cdim[0] * (uint32_t)w.cdim[1]));
    b[0] = (cx * 4 + (int32_t)(tb & 3u)) * 4 + (int32_t)(bb & 3u);
    b[1] = (cy * 4 + (int32_t)(tb >> 4)) * 4 + (int32_t)(bb >> 4);
    b[2] = (cz * 4 + (int32_t)((tb >> 2) & 3u)) * 4 + (int32_t)((bb >> 2) & 3u);
}

// Slot brick coordinates -> nested brick index.
__device__ __forceinline__ uint32_t brick_nested(const DevWorld &w, const int32_t b[3]) {
    const uint32_t ci = (uint32_t)(b[0] >> 4) + (uint32_t)(b[1] >> 4) * (uint32_t)w.cdim[0] +
                        (uint32_t)(b[2] >> 4) * (uint32_t)w.cdim[0] * (uint32_t)w.cdim[1];
    const uint32_t tn = ci * 64u + (uint32_t)cbit((b[0] >> 2) & 3, (b[1] >> 2) & 3, (b[2] >> 2) & 3);
    return tn * 64u + (uint32_t)cbit(b[0] & 3, b[1] & 3, b[2] & 3);
}

// World-brick range of an edit clipped to the window, and the sphere / brick-box test.
struct EditRange {
    int32_t lo[3], hi[3];
};
__host__ __device__ __forceinline__ bool edit_range(const vg_edit &e, const int32_t origin[3],
                                                    const int32_t dims[3], EditRange &r) {
    for (int k = 0; k < 3; ++k) {
        int32_t a = e.center[k] - (int32_t)e.radius, b = e.center[k] + (int32_t)e.radius;
        a = a > origin[k] ? a : origin[k];
        b = b < origin[k] + dims[k] - 1 ? b : origin[k] + dims[k] - 1;
        if (a > b) return false;
        r.lo[k] = a >> 2; // the origin is chunk-aligned: world bricks are slot bricks
        r.hi[k] = b >> 2;
    }
    return true;
}
__host__ __device__ __forceinline__ bool sphere_touches_brick(const vg_edit &e, const int32_t wb[3]) {
    int64_t d2 = 0;
    for (int k = 0; k < 3; ++k) {
        const int32_t lo = wb[k] * 4, hi = lo + 3, c = e.center[k];
        const int32_t d = c < lo ? lo - c : (c > hi ? c - hi : 0);
        d2 += (int64_t)d * d;
    }
    return d2 <= (int64_t)e.radius * e.radius;
}

// ------------------------------------------------------------------ kernels --
__global__ void k_edit_clear(uint32_t *a, uint32_t na, uint32_t *b, uint32_t nb, uint32_t *c,
                             uint32_t nc, uint32_t *tile_n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0u;
    if (i < nb) b[i] = 0u;
    if (i < nc) c[i] = 0u;
    if (i == 0) *tile_n = 0u;
}

// One wave per edit, one lane per brick of its clipped brick box. COUNT: per-brick pair
// counters; the first pair of a brick flags its tile and the first flag appends the tile to
// the touched list. FILL: each pair takes a slot of its brick's run
// [end[bn] - cnt[bn], end[bn]) and counts the counter back down (zero again afterwards).
template <bool FILL>
__global__ __launch_bounds__(256) void k_edit_enum(DevWorld w, const vg_edit *__restrict__ edits,
                                                   uint32_t n, uint32_t *__restrict__ cnt,
                                                   const uint32_t *__restrict__ end,
                                                   uint32_t *__restrict__ list,
                                                   uint32_t *__restrict__ tile_flag,
                                                   uint32_t *__restrict__ tile_list,
                                                   uint32_t *__restrict__ tile_n) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (i >= n) return; // wave-uniform
    const vg_edit e = edits[i];
    EditRange r;
    if (!edit_range(e, w.origin, w.dims, r)) return;
    const uint32_t nx = (uint32_t)(r.hi[0] - r.lo[0] + 1), nz = (uint32_t)(r.hi[2] - r.lo[2] + 1);
    const uint32_t vol = nx * nz * (uint32_t)(r.hi[1] - r.lo[1] + 1);
    for (uint32_t j = lane; j < vol; j += 64u) {
        const uint32_t t = j / nx;
        const int32_t wb[3] = {r.lo[0] + (int32_t)(j - t * nx), r.lo[1] + (int32_t)(t / nz),
                               r.lo[2] + (int32_t)(t % nz)};
        if (!sphere_touches_brick(e, wb)) continue;
        int32_t sb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int32_t sv = (wb[k] * 4) % w.dims[k];
            sv = sv < 0 ? sv + w.dims[k] : sv;
            sb[k] = sv >> 2;
        }
        const uint32_t bn = brick_nested(w, sb);
        if (FILL) {
            const uint32_t c = atomicSub(cnt + bn, 1u);
            list[end[bn] - c] = i;
        } else if (atomicAdd(cnt + bn, 1u) == 0u) {
            const uint32_t tn = bn >> 6;
            if (atomicExch(tile_flag + tn, 1u) == 0u) tile_list[atomicAdd(tile_n, 1u)] = tn;
        }
    }
}

__device__ __forceinline__ uint32_t lane_read(uint32_t v, uint32_t k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)k);
}

// Blocks of 16 waves walk the touched-tile list; in a tile, wave k takes the touched bricks
// k, k + 16, ... (in brick order), one lane per voxel. Each touched brick applies its edits: a
// voxel some sphere covers takes the material of the covering edit of the highest index
// with Brick::set semantics (src/voxel.rs:315-325), the others keep their bit and byte. The
// tile's flag is cleared for the next batch.
constexpr uint32_t kBrickWaves = 16;
__global__ __launch_bounds__(1024) void k_edit_bricks(DevWorld w, const vg_edit *__restrict__ edits,
                                                      const uint32_t *__restrict__ end,
                                                      const uint32_t *__restrict__ list,
                                                      const uint32_t *__restrict__ tile_list,
                                                      const uint32_t *__restrict__ tile_n,
                                                      uint32_t *__restrict__ tile_flag,
                                                      uint64_t *__restrict__ occ,
                                                      uint8_t *__restrict__ mat,
                                                      uint8_t *__restrict__ bu,
                                                      uint32_t *__restrict__ tile_dirty) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t nt = *tile_n;
    // voxel of this lane inside a brick: bit x + 4z + 16y
    const int32_t vx = (int32_t)(lane & 3u), vy = (int32_t)(lane >> 4), vz = (int32_t)((lane >> 2) & 3u);
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const uint32_t tn = tile_list[t];
        const uint32_t bl = tn * 64u + lane;
        const uint32_t hi_l = end[bl];
        uint32_t lo_l = __shfl_up(hi_l, 1, 64);
        if (lane == 0) lo_l = bl ? end[bl - 1u] : 0u;
        uint64_t todo = __ballot(hi_l > lo_l);
        for (uint32_t k = 0; k < wave && todo; ++k) todo &= todo - 1; // skip to this wave's first
        bool tile_changed = false;
        while (todo) {
            const uint32_t src = (uint32_t)(__ffsll((long long)todo) - 1);
            for (uint32_t k = 0; k < kBrickWaves && todo; ++k) todo &= todo - 1;
            const uint32_t bn = tn * 64u + src;
            const uint32_t lo = lane_read(lo_l, src), hi = lane_read(hi_l, src);
            int32_t sb[3];
            brick_slot(w, bn, sb);
            const int32_t wv[3] = {slot_to_world(sb[0] * 4 + vx, w.origin[0], w.dims[0]),
                                   slot_to_world(sb[1] * 4 + vy, w.origin[1], w.dims[1]),
                                   slot_to_world(sb[2] * 4 + vz, w.origin[2], w.dims[2])};
            const uint64_t old_occ = occ[bn];
            const bool old_bit = (old_occ >> lane) & 1u;
            const uint32_t raw = mat[(size_t)bn * 64u + lane];
            int32_t best = -1;
            uint32_t m = raw;
            for (uint32_t base = lo; base < hi; base += 64u) {
                const uint32_t kn = hi - base < 64u ? hi - base : 64u;
                uint32_t ei = 0, ex = 0, ey = 0, ez = 0, er = 0;
                if (lane < kn) {
                    ei = list[base + lane];
                    const vg_edit e = edits[ei];
                    ex = (uint32_t)e.center[0];
                    ey = (uint32_t)e.center[1];
                    ez = (uint32_t)e.center[2];
                    er = ((uint32_t)e.radius * (uint32_t)e.radius) << 8 | (uint32_t)e.material;
                }
                for (uint32_t k = 0; k < kn; ++k) {
                    const int32_t idx = (int32_t)lane_read(ei, k);
                    const uint32_t rm = lane_read(er, k);
                    const int32_t dx = wv[0] - (int32_t)lane_read(ex, k);
                    const int32_t dy = wv[1] - (int32_t)lane_read(ey, k);
                    const int32_t dz = wv[2] - (int32_t)lane_read(ez, k);
                    if (dx * dx + dy * dy + dz * dz <= (int32_t)(rm >> 8) && idx > best) {
                        best = idx;
                        m = rm & 0xFFu;
                    }
                }
            }
            const bool touched = best >= 0;
            const bool bit = touched ? m != 0u : old_bit;
            const uint64_t new_occ = __ballot(bit);
            tile_changed |= __ballot(touched && (m != raw || bit != old_bit)) != 0;
            const uint32_t m0 = __shfl(m, 0, 64);
            const bool all_same = __ballot(m == m0) == ~0ull;
            if (touched) mat[(size_t)bn * 64u + lane] = (uint8_t)m;
            if (lane == 0) {
                occ[bn] = new_occ;
                bu[bn] = (new_occ == ~0ull && all_same && m0 != 0u) ? (uint8_t)m0 : (uint8_t)0;
            }
        }
        if (lane == 0) {
            if (tile_changed) atomicOr(tile_dirty + (tn >> 5), 1u << (tn & 31u));
            if (wave == 0) tile_flag[tn] = 0u;
        }
    }
}

// 1 + the highest window-local voxel row a tile can hold (0: an empty tile): nested tile tn
// (chunk tn / 64 in the reference order, tile bit tn % 64), its word, whether it is uniform.
// Its row in a window that folds y is a slot row, but such a window's bound is dims[1].
__device__ __forceinline__ int32_t tile_ytop(uint32_t tn, uint64_t word, bool uni, uint32_t cdim0,
                                             uint32_t cdim1) {
    const uint32_t ci = tn >> 6;
    const int32_t ty = (int32_t)(((ci / cdim0) % cdim1) * 4u + ((tn & 63u) >> 4));
    if (uni) return ty * 16 + 16;
    if (!word) return 0;
    return ty * 16 + ((63 - __clzll((long long)word)) >> 4) * 4 + 4;
}

// Raise the world's y bound to `top` (it only grows; most tiles leave it as it is).
__device__ __forceinline__ void raise_ytop(int32_t *ytop, int32_t top) {
    if (top > __hip_atomic_load(ytop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(ytop, top);
}

// Waves walk the touched-tile list: a dirty tile recomputes its word and uniform byte from
// its 64 bricks, marks its chunk and raises the world's y bound (never lowered by an edit).
__global__ __launch_bounds__(256) void k_edit_tiles(const uint32_t *__restrict__ tile_dirty,
                                                    const uint32_t *__restrict__ tile_list,
                                                    const uint32_t *__restrict__ tile_n,
                                                    const uint64_t *__restrict__ occ,
                                                    const uint8_t *__restrict__ bu,
                                                    uint64_t *__restrict__ tile,
                                                    uint8_t *__restrict__ tu,
                                                    uint32_t *__restrict__ chunk_dirty,
                                                    uint32_t cdim0, uint32_t cdim1, int32_t *ytop) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nt = *tile_n;
    for (uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6); t < nt; t += gridDim.x * 4u) {
        const uint32_t tn = tile_list[t];
        if (!((tile_dirty[tn >> 5] >> (tn & 31u)) & 1u)) continue; // wave-uniform
        const uint32_t bn = tn * 64u + lane;
        const uint64_t word = __ballot(occ[bn] != 0ull);
        const uint32_t b = bu[bn];
        const uint32_t b0 = __shfl(b, 0, 64);
        const bool uni = b0 != 0u && __ballot(b == b0) == ~0ull;
        if (lane == 0) {
            const uint64_t prev = tile[tn];
            tile[tn] = word;
            tu[tn] = uni ? (uint8_t)b0 : (uint8_t)0;
            if ((prev == 0ull) != (word == 0ull)) {
                const uint32_t ci = tn >> 6;
                atomicOr(chunk_dirty + (ci >> 5), 1u << (ci & 31u));
            }
            raise_ytop(ytop, tile_ytop(tn, word, uni, cdim0, cdim1));
        }
    }
}

__global__ __launch_bounds__(256) void k_edit_chunks(const uint32_t *__restrict__ chunk_dirty,
                                                     DevWorld w, uint32_t n_chunks,
                                                     const uint64_t *__restrict__ tile,
                                                     uint64_t *__restrict__ chunk,
                                                     uint32_t *__restrict__ l4_dirty) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t ci = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (ci >= n_chunks || !((chunk_dirty[ci >> 5] >> (ci & 31u)) & 1u)) return; // wave-uniform
    const uint64_t word = __ballot(tile[ci * 64u + lane] != 0ull);
    if (lane == 0) {
        const uint64_t prev = chunk[ci];
        chunk[ci] = word;
        if ((prev == 0ull) != (word == 0ull)) {
            const uint32_t cx = ci % (uint32_t)w.cdim[0];
            const uint32_t cy = (ci / (uint32_t)w.cdim[0]) % (uint32_t)w.cdim[1];
            const uint32_t cz = ci / ((uint32_t)w.cdim[0] * (uint32_t)w.cdim[1]);
            const uint32_t li = (cx >> 2) + (cy >> 2) * (uint32_t)w.l4dim[0] +
                                (cz >> 2) * (uint32_t)w.l4dim[0] * (uint32_t)w.l4dim[1];
            atomicOr(l4_dirty + (li >> 5), 1u << (li & 31u));
        }
    }
}

__global__ void k_edit_l4(const uint32_t *__restrict__ l4_dirty, DevWorld w, uint32_t n_l4,
                          const uint64_t *__restrict__ chunk, uint64_t *__restrict__ l4) {
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n_l4 || !((l4_dirty[li >> 5] >> (li & 31u)) & 1u)) return;
    const uint32_t lx = li % (uint32_t)w.l4dim[0];
    const uint32_t ly = (li / (uint32_t)w.l4dim[0]) % (uint32_t)w.l4dim[1];
    const uint32_t lz = li / ((uint32_t)w.l4dim[0] * (uint32_t)w.l4dim[1]);
    uint64_t word = 0;
    for (int b = 0; b < 64; ++b) {
        const uint32_t cx = lx * 4u + (uint32_t)(b & 3), cy = ly * 4u + (uint32_t)(b >> 4),
                       cz = lz * 4u + (uint32_t)((b >> 2) & 3);
        if (cx >= (uint32_t)w.cdim[0] || cy >= (uint32_t)w.cdim[1] || cz >= (uint32_t)w.cdim[2]) continue;
        const uint32_t ci = cx + cy * (uint32_t)w.cdim[0] + cz * (uint32_t)w.cdim[0] * (uint32_t)w.cdim[1];
        if (chunk[ci] != 0ull) word |= 1ull << b;
    }
    l4[li] = word;
}

// --------------------------------------------------------------- projection --
struct ProjCam {
    float o[3], f[3], r[3], u[3];
    float tan_half, aspect, w, h, jx, jy;
};

// Screen rect (pixels, inclusive) of the convex hull of n points; 0 = nothing visible,
// 1 = rect, 2 = the whole screen (a point at or behind the camera plane while another is
// in front).
__device__ __forceinline__ int project_points(const ProjCam &c, const float (*p)[3], int n,
                                              float rect[4]) {
    int front = 0, back = 0;
    float x0 = 1e30f, y0 = 1e30f, x1 = -1e30f, y1 = -1e30f;
    for (int i = 0; i < n; ++i) {
        float d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] = p[i][k] - c.o[k];
        const float z = d[0] * c.f[0] + d[1] * c.f[1] + d[2] * c.f[2];
        if (z <= 1e-3f) {
            ++back;
            continue;
        }
        ++front;
        const float x = d[0] * c.r[0] + d[1] * c.r[1] + d[2] * c.r[2];
        const float y = d[0] * c.u[0] + d[1] * c.u[1] + d[2] * c.u[2];
        const float nx = x / (z * c.tan_half * c.aspect), ny = y / (z * c.tan_half);
        const float px = (nx + 1.0f) * 0.5f * c.w - 0.5f - c.jx;
        const float py = (1.0f - ny) * 0.5f * c.h - 0.5f - c.jy;
        x0 = fminf(x0, px);
        x1 = fmaxf(x1, px);
        y0 = fminf(y0, py);
        y1 = fmaxf(y1, py);
    }
    if (front == 0) return 0;
    if (back > 0) return 2;
    rect[0] = x0 - 2.0f;
    rect[1] = y0 - 2.0f;
    rect[2] = x1 + 2.0f;
    rect[3] = y1 + 2.0f;
    return 1;
}

// Mark the 8x8 blocks a pixel rect touches; the wave's lanes share the (block row, mask
// word) pairs of the rect, one atomic per pair.
__device__ __forceinline__ void mark_rect(uint32_t *mask, uint32_t nbx, uint32_t nby, int32_t width,
                                          int32_t height, const float rect[4], uint32_t lane) {
    const float xa = fmaxf(rect[0], 0.0f), ya = fmaxf(rect[1], 0.0f);
    const float xb = fminf(rect[2], (float)(width - 1)), yb = fminf(rect[3], (float)(height - 1));
    if (xa > xb || ya > yb) return;
    const uint32_t bx0 = (uint32_t)xa / 8u, bx1 = min((uint32_t)xb / 8u, nbx - 1u);
    const uint32_t by0 = (uint32_t)ya / 8u, by1 = min((uint32_t)yb / 8u, nby - 1u);
    const uint32_t span = (bx1 - bx0) / 32u + 2u; // words a row of the rect can touch
    const uint32_t total = (by1 - by0 + 1u) * span;
    for (uint32_t i = lane; i < total; i += 64u) {
        const uint32_t by = by0 + i / span, j = i % span;
        const uint32_t first_b = by * nbx + bx0, last_b = by * nbx + bx1;
        const uint32_t word = (first_b >> 5) + j;
        if (word > (last_b >> 5)) continue;
        const uint32_t lo = word == (first_b >> 5) ? (first_b & 31u) : 0u;
        const uint32_t hi = word == (last_b >> 5) ? (last_b & 31u) : 31u;
        const uint32_t bits = (hi == 31u ? ~0u : ((1u << (hi + 1u)) - 1u)) & ~((1u << lo) - 1u);
        atomicOr(mask + word, bits);
    }
}

// One wave's share of the dirty-block mask for nested tile tn: mark the 8x8 blocks whose
// primary rays can enter the tile box, and (shadow set) whose primary hits can cast a shadow
// ray through it. mask[nwords] is the whole-screen flag.
__device__ void tile_blocks(const DevWorld &w, uint32_t tn, const ProjCam &c, int32_t width,
                            int32_t height, uint32_t nbx, uint32_t nby, int shadow, float sx,
                            float sy, float sz, float stmax, uint32_t *mask, uint32_t nwords,
                            uint32_t lane) {
    int32_t sb[3];
    brick_slot(w, tn * 64u, sb); // the tile's first brick
    float lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int32_t wv = slot_to_world(sb[k] * 4, w.origin[k], w.dims[k]);
        lo[k] = (float)wv - 0.01f;
        hi[k] = (float)(wv + 16) + 0.01f;
    }
    float pts[16][3];
    for (int i = 0; i < 8; ++i) {
        pts[i][0] = (i & 1) ? hi[0] : lo[0];
        pts[i][1] = (i & 2) ? hi[1] : lo[1];
        pts[i][2] = (i & 4) ? hi[2] : lo[2];
    }
    float rect[4];
    int r = project_points(c, pts, 8, rect);
    if (r == 2) {
        if (lane == 0) mask[nwords] = 1u;
        return;
    }
    if (r == 1) mark_rect(mask, nbx, nby, width, height, rect, lane);
    if (!shadow) return;
    // hit points p with p + s t inside the box for some t in [0, stmax]: the box swept by
    // -s t, t clipped to where the swept box still meets the window
    const float s[3] = {sx, sy, sz};
    float tmax = stmax;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float wlo = (float)w.origin[k], whi = (float)(w.origin[k] + w.dims[k]);
        if (s[k] > 1e-6f) tmax = fminf(tmax, (hi[k] - wlo) / s[k]);
        else if (s[k] < -1e-6f) tmax = fminf(tmax, (whi - lo[k]) / -s[k]);
    }
    tmax = fmaxf(tmax, 0.0f) + 0.5f;
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) pts[8 + i][k] = pts[i][k] - s[k] * tmax;
    r = project_points(c, pts, 16, rect);
    if (r == 2) {
        if (lane == 0) mask[nwords] = 1u;
        return;
    }
    if (r == 1) mark_rect(mask, nbx, nby, width, height, rect, lane);
}

// Waves walk the touched-tile list of the last edit batch and mark the blocks of its changed
// tiles; once the whole-screen flag is set the remaining tiles have nothing to add.
__global__ __launch_bounds__(256) void k_edit_blocks(DevWorld w, const uint32_t *__restrict__ tile_dirty,
                                                     const uint32_t *__restrict__ tile_list,
                                                     const uint32_t *__restrict__ tile_n, ProjCam c,
                                                     int32_t width, int32_t height, uint32_t nbx,
                                                     uint32_t nby, int shadow, float sx, float sy,
                                                     float sz, float stmax, uint32_t *mask,
                                                     uint32_t nwords) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nt = *tile_n;
    for (uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6); t < nt; t += gridDim.x * 4u) {
        const uint32_t tn = tile_list[t];
        if (!((tile_dirty[tn >> 5] >> (tn & 31u)) & 1u)) continue; // wave-uniform
        if (__shfl(__atomic_load_n(mask + nwords, __ATOMIC_RELAXED), 0, 64) != 0u) return;
        tile_blocks(w, tn, c, width, height, nbx, nby, shadow, sx, sy, sz, stmax, mask, nwords, lane);
    }
}

__global__ void k_mask_clear(uint32_t *mask, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mask[i] = 0u;
}

// Compact a block mask (whole-screen flag at mask[nwords]) into a list of block ids; one
// atomic per wave.
__global__ __launch_bounds__(256) void k_mask_compact(const uint32_t *__restrict__ mask,
                                                      uint32_t nwords, uint32_t nblocks,
                                                      uint32_t *__restrict__ list,
                                                      uint32_t *list_n) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    const bool all = mask[nwords] != 0u;
    const bool on = b < nblocks && (all || ((mask[b >> 5] >> (b & 31u)) & 1u));
    const uint64_t m = __ballot(on);
    if (!m) return; // wave-uniform
    const uint32_t lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(list_n, (uint32_t)__popcll(m));
    base = __shfl(base, 0, 64);
    if (on) list[base + lanes_below(m)] = b;
}

__global__ void k_list_reset(uint32_t *list_n) { *list_n = 0u; }

// ---------------------------------------------------------------- download --
__global__ void k_download_bricks(const uint64_t *__restrict__ occ, const uint8_t *__restrict__ mat,
                                  const uint8_t *__restrict__ bu, int64_t first, int64_t count,
                                  int32_t bx_n, int32_t by_n, int32_t cx_n, int32_t cy_n,
                                  vg_brick *__restrict__ out, uint8_t *__restrict__ bu_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t bi = first + i;
    const int32_t bx = (int32_t)(bi % bx_n), by = (int32_t)((bi / bx_n) % by_n),
                  bz = (int32_t)(bi / ((int64_t)bx_n * by_n));
    const int64_t ci = (bx >> 4) + (int64_t)(by >> 4) * cx_n + (int64_t)(bz >> 4) * cx_n * cy_n;
    const int64_t tn = ci * 64 + cbit((bx >> 2) & 3, (by >> 2) & 3, (bz >> 2) & 3);
    const int64_t bn = tn * 64 + cbit(bx & 3, by & 3, bz & 3);
    uint64_t *dst = reinterpret_cast<uint64_t *>(out + i);
    dst[0] = occ[bn];
    const uint64_t *src = reinterpret_cast<const uint64_t *>(mat + bn * 64);
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[1 + k] = src[k];
    bu_out[i] = bu[bn];
}

__global__ void k_download_tiles(const uint64_t *__restrict__ tile, const uint8_t *__restrict__ tu,
                                 int64_t n, int32_t tx_n, int32_t ty_n, int32_t cx_n, int32_t cy_n,
                                 uint64_t *__restrict__ tile_out, uint8_t *__restrict__ tu_out) {
    const int64_t ti = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= n) return;
    const int32_t tx = (int32_t)(ti % tx_n), ty = (int32_t)((ti / tx_n) % ty_n),
                  tz = (int32_t)(ti / ((int64_t)tx_n * ty_n));
    const int64_t ci = (tx >> 2) + (int64_t)(ty >> 2) * cx_n + (int64_t)(tz >> 2) * cx_n * cy_n;
    const int64_t tn = ci * 64 + cbit(tx & 3, ty & 3, tz & 3);
    tile_out[ti] = tile[tn];
    tu_out[ti] = tu[tn];
}

// ------------------------------------------------------------- generation --
// The benchmark terrain generated on the device: the same bytes as vg_generate_terrain +
// vg_build_pyramid (src/voxel.rs:1163-1232 noise and heights, the generator's column rule),
// fp32 in the host's operation order (value noise with smoothstep weights, 5-octave fBm,
// u32 wrapping hashes; no contraction).
__device__ __forceinline__ float gen_hash2(int32_t x, int32_t z) {
    uint32_t h = (uint32_t)x * 0x9E3779B1u + (uint32_t)z * 0x85EBCA77u;
    h *= 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    return (float)(h & 0xFFFFFFu) / (float)0xFFFFFF * 2.0f - 1.0f;
}
__device__ __forceinline__ float gen_hash3(int32_t x, int32_t y, int32_t z) {
    uint32_t h = (uint32_t)x * 0x9E3779B1u + (uint32_t)y * 0x85EBCA77u + (uint32_t)z * 0xC2B2AE3Du;
    h *= 0xD2B74407u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    return (float)(h & 0xFFFFFFu) / (float)0xFFFFFF * 2.0f - 1.0f;
}
__device__ __forceinline__ float gen_smooth(float t) { return t * t * (3.0f - 2.0f * t); }
__device__ __forceinline__ float gen_lerp(float a, float b, float w) { return a * (1.0f - w) + b * w; }
__device__ __forceinline__ float gen_noise2(float x, float z) {
    const int32_t xi = (int32_t)floorf(x), zi = (int32_t)floorf(z);
    const float xf = gen_smooth(x - (float)xi), zf = gen_smooth(z - (float)zi);
    const float a = gen_lerp(gen_hash2(xi, zi), gen_hash2(xi + 1, zi), xf);
    const float b = gen_lerp(gen_hash2(xi, zi + 1), gen_hash2(xi + 1, zi + 1), xf);
    return gen_lerp(a, b, zf);
}
__device__ __forceinline__ float gen_noise3(float x, float y, float z) {
    const int32_t xi = (int32_t)floorf(x), yi = (int32_t)floorf(y), zi = (int32_t)floorf(z);
    const float xf = gen_smooth(x - (float)xi), yf = gen_smooth(y - (float)yi),
                zf = gen_smooth(z - (float)zi);
    const float a = gen_lerp(gen_hash3(xi, yi, zi), gen_hash3(xi + 1, yi, zi), xf);
    const float b = gen_lerp(gen_hash3(xi, yi + 1, zi), gen_hash3(xi + 1, yi + 1, zi), xf);
    const float c = gen_lerp(gen_hash3(xi, yi, zi + 1), gen_hash3(xi + 1, yi, zi + 1), xf);
    const float d = gen_lerp(gen_hash3(xi, yi + 1, zi + 1), gen_hash3(xi + 1, yi + 1, zi + 1), xf);
    return gen_lerp(gen_lerp(a, b, yf), gen_lerp(c, d, yf), zf);
}
__device__ __forceinline__ int32_t gen_height(float sx, float sz, float top, int32_t x, int32_t z) {
    float total = 0.0f, amp = 1.0f, freq = 1.0f, max_amp = 0.0f;
    const float fx = ((float)x + sx) * 0.007f, fz = ((float)z + sz) * 0.007f;
    for (int i = 0; i < 5; ++i) {
        total += gen_noise2(fx * freq, fz * freq) * amp;
        max_amp += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    float h = 64.0f + 40.0f * (total / max_amp);
    h = fminf(fmaxf(h, 2.0f), top);
    return (int32_t)h;
}

// One wave per slot brick column (bx, bz), one lane per voxel of a brick: every brick of the
// column from y = 0 up is written whole (occupancy from the ballot, the material bytes, the
// uniform byte).
__global__ __launch_bounds__(256) void k_gen_bricks(DevWorld w, float sx, float sz, uint32_t ncols,
                                                    uint64_t *__restrict__ occ,
                                                    uint8_t *__restrict__ mat,
                                                    uint8_t *__restrict__ bu) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t col = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (col >= ncols) return; // wave-uniform
    const int32_t nbx = w.dims[0] >> 2;
    const int32_t bx = (int32_t)(col % (uint32_t)nbx), bz = (int32_t)(col / (uint32_t)nbx);
    const int32_t cx = (int32_t)(lane & 3u), cz = (int32_t)((lane >> 2) & 3u), vy = (int32_t)(lane >> 4);
    const int32_t wx = slot_to_world(bx * 4 + cx, w.origin[0], w.dims[0]);
    const int32_t wz = slot_to_world(bz * 4 + cz, w.origin[2], w.dims[2]);
    const int32_t Y = w.dims[1];
    const int32_t h = gen_height(sx, sz, (float)(Y - 1), wx, wz);
    const int32_t ytop = min(max(h, 55), Y - 1);
    for (int32_t by = 0; by < (Y >> 2); ++by) {
        const int32_t y = by * 4 + vy;
        uint32_t m = 0;
        if (y <= ytop) {
            if (y > h) {
                m = 12u; // water L8 in open air below y = 56
            } else if (!(y > 4 && y + 3 < h &&
                         gen_noise3((float)wx * 0.045f, (float)y * 0.085f, (float)wz * 0.045f) > 0.6f)) {
                m = y == h ? (h <= 57 ? 1u : 2u) : (y + 4 >= h ? 3u : 4u);
            }
        }
        const int32_t sb[3] = {bx, by, bz};
        const uint32_t bn = brick_nested(w, sb);
        const uint64_t o = __ballot(m != 0u);
        const uint32_t m0 = __shfl(m, 0, 64);
        const bool same = __ballot(m == m0) == ~0ull;
        mat[(size_t)bn * 64u + lane] = (uint8_t)m;
        if (lane == 0) {
            occ[bn] = o;
            bu[bn] = (o == ~0ull && same && m0 != 0u) ? (uint8_t)m0 : (uint8_t)0;
        }
    }
}

// The pyramid of a whole world, bottom up: one wave per nested tile (word = ballot of its
// non-empty bricks, uniform byte when all 64 bricks share one uniform material), one wave
// per chunk, one thread per L4 cell.
__global__ __launch_bounds__(256) void k_gen_tiles(uint32_t n_tiles, const uint64_t *__restrict__ occ,
                                                   const uint8_t *__restrict__ bu,
                                                   uint64_t *__restrict__ tile, uint8_t *__restrict__ tu,
                                                   uint32_t cdim0, uint32_t cdim1, int32_t *ytop) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tn = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (tn >= n_tiles) return; // wave-uniform
    const uint32_t bn = tn * 64u + lane;
    const uint64_t word = __ballot(occ[bn] != 0ull);
    const uint32_t b = bu[bn];
    const uint32_t b0 = __shfl(b, 0, 64);
    const bool uni = b0 != 0u && __ballot(b == b0) == ~0ull;
    if (lane == 0) {
        tile[tn] = word;
        tu[tn] = uni ? (uint8_t)b0 : (uint8_t)0;
        raise_ytop(ytop, tile_ytop(tn, word, uni, cdim0, cdim1));
    }
}

// The generated world's y bound starts at 0 (dims[1] for a window that folds y); the tile
// pass raises it.
__global__ void k_set_ytop(int32_t *ytop, int32_t v) { *ytop = v; }

__global__ __launch_bounds__(256) void k_gen_chunks(uint32_t n_chunks, const uint64_t *__restrict__ tile,
                                                    uint64_t *__restrict__ chunk) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t ci = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (ci >= n_chunks) return; // wave-uniform
    const uint64_t word = __ballot(tile[ci * 64u + lane] != 0ull);
    if (lane == 0) chunk[ci] = word;
}

__global__ void k_gen_l4(DevWorld w, uint32_t n_l4, const uint64_t *__restrict__ chunk,
                         uint64_t *__restrict__ l4) {
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= n_l4) return;
    const uint32_t lx = li % (uint32_t)w.l4dim[0];
    const uint32_t ly = (li / (uint32_t)w.l4dim[0]) % (uint32_t)w.l4dim[1];
    const uint32_t lz = li / ((uint32_t)w.l4dim[0] * (uint32_t)w.l4dim[1]);
    uint64_t word = 0;
    for (int b = 0; b < 64; ++b) {
        const uint32_t cx = lx * 4u + (uint32_t)(b & 3), cy = ly * 4u + (uint32_t)(b >> 4),
                       cz = lz * 4u + (uint32_t)((b >> 2) & 3);
        if (cx >= (uint32_t)w.cdim[0] || cy >= (uint32_t)w.cdim[1] || cz >= (uint32_t)w.cdim[2]) continue;
        const uint32_t ci = cx + cy * (uint32_t)w.cdim[0] + cz * (uint32_t)w.cdim[0] * (uint32_t)w.cdim[1];
        if (chunk[ci] != 0ull) word |= 1ull << b;
    }
    l4[li] = word;
}

// --------------------------------------------------------------------- host --
static void free_edit_state(vg_edit_state *s) {
    void *dev[] = {s->edits, s->cnt, s->end, s->list, s->temp, s->tile_flag, s->tile_list,
                   s->tile_n, s->tile_dirty, s->chunk_dirty, s->l4_dirty, s->blist, s->blist_n};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    for (int i = 0; i < 2; ++i) {
        if (s->staging[i]) (void)hipHostFree(s->staging[i]);
        if (s->staged[i]) (void)hipEventDestroy(s->staged[i]);
    }
    delete s;
}

static int edit_state(vg_world *w, vg_edit_state **out) {
    if (!w->edit) {
        vg_edit_state *s = new vg_edit_state();
        std::memset(s, 0, sizeof(*s));
        const Dims &d = w->d;
        const size_t tw = (size_t)(d.n_tiles + 31) / 32, cw = (size_t)(d.n_chunks + 31) / 32,
                     lw = (size_t)(d.n_l4 + 31) / 32;
        bool ok = hipMalloc((void **)&s->tile_flag, (size_t)d.n_tiles * 4) == hipSuccess &&
                  hipMalloc((void **)&s->tile_list, (size_t)d.n_tiles * 4) == hipSuccess &&
                  hipMalloc((void **)&s->tile_n, 256) == hipSuccess &&
                  hipMalloc((void **)&s->tile_dirty, tw * 4) == hipSuccess &&
                  hipMalloc((void **)&s->chunk_dirty, cw * 4) == hipSuccess &&
                  hipMalloc((void **)&s->l4_dirty, lw * 4) == hipSuccess &&
                  hipMalloc((void **)&s->blist_n, 256) == hipSuccess &&
                  hipMalloc((void **)&s->cnt, (size_t)d.n_bricks * 4) == hipSuccess &&
                  hipMalloc((void **)&s->end, (size_t)d.n_bricks * 4) == hipSuccess;
        if (ok) ok = hipMemset(s->tile_flag, 0, (size_t)d.n_tiles * 4) == hipSuccess &&
                     hipMemset(s->tile_n, 0, 256) == hipSuccess &&
                     hipMemset(s->tile_dirty, 0, tw * 4) == hipSuccess &&
                     hipMemset(s->chunk_dirty, 0, cw * 4) == hipSuccess &&
                     hipMemset(s->l4_dirty, 0, lw * 4) == hipSuccess &&
                     hipMemset(s->cnt, 0, (size_t)d.n_bricks * 4) == hipSuccess;
        for (int i = 0; i < 2 && ok; ++i)
            ok = hipEventCreateWithFlags(&s->staged[i], hipEventDisableTiming) == hipSuccess;
        const size_t tb = (size_t)scan_scratch_words((uint32_t)d.n_bricks) * sizeof(uint32_t);
        if (ok) ok = hipMalloc(&s->temp, tb) == hipSuccess;
        if (!ok) {
            free_edit_state(s);
            return fail(VG_E_NOMEM, "edit state");
        }
        // the counters must be zero before the first batch's kernels run on any stream
        if (hipDeviceSynchronize() != hipSuccess) {
            free_edit_state(s);
            return fail(VG_E_DEVICE, "edit state");
        }
        s->temp_bytes = tb;
        w->edit = s;
    }
    *out = w->edit;
    return VG_OK;
}

static int ensure_edit_capacity(vg_edit_state *s, size_t n_edits, size_t n_pairs) {
    if (n_edits > s->cap_edits) {
        HIP_TRY(hipDeviceSynchronize());
        if (s->edits) (void)hipFree(s->edits);
        s->edits = nullptr;
        for (int i = 0; i < 2; ++i) {
            if (s->staging[i]) (void)hipHostFree(s->staging[i]);
            s->staging[i] = nullptr;
            s->staged_rec[i] = false;
        }
        s->cap_edits = 0;
        const size_t cap = std::max<size_t>(n_edits, 4096);
        if (hipMalloc((void **)&s->edits, cap * sizeof(vg_edit)) != hipSuccess ||
            hipHostMalloc((void **)&s->staging[0], cap * sizeof(vg_edit)) != hipSuccess ||
            hipHostMalloc((void **)&s->staging[1], cap * sizeof(vg_edit)) != hipSuccess)
            return fail(VG_E_NOMEM, "edit buffers for %zu edits", cap);
        s->cap_edits = cap;
    }
    if (n_pairs > s->cap_pairs) {
        HIP_TRY(hipDeviceSynchronize());
        if (s->list) (void)hipFree(s->list);
        s->list = nullptr;
        s->cap_pairs = 0;
        const size_t cap = std::max<size_t>(n_pairs + n_pairs / 4, 1u << 16);
        if (hipMalloc((void **)&s->list, cap * 4) != hipSuccess)
            return fail(VG_E_NOMEM, "edit pairs (%zu)", cap);
        s->cap_pairs = cap;
    }
    return VG_OK;
}

static ProjCam proj_cam(const vg_camera *c, int32_t width, int32_t height) {
    ProjCam p;
    for (int k = 0; k < 3; ++k) {
        p.o[k] = c->origin[k];
        p.f[k] = c->forward[k];
        p.r[k] = c->right[k];
        p.u[k] = c->up[k];
    }
    p.tan_half = c->tan_half_fov;
    p.w = c->resolution[0] > 0.0f ? c->resolution[0] : (float)width;
    p.h = c->resolution[1] > 0.0f ? c->resolution[1] : (float)height;
    p.aspect = p.w / p.h;
    p.jx = c->jitter[0];
    p.jy = c->jitter[1];
    return p;
}

} // namespace vg

using namespace vg;

void vg_edit_state_free(vg_world *w) {
    if (w && w->edit) {
        free_edit_state(w->edit);
        w->edit = nullptr;
    }
}

extern "C" {

int vg_world_apply_edits(vg_world *w, const vg_edit *edits, uint32_t n, void *stream) {
    int st = check_world(w);
    if (st) return st;
    if (n && !edits) return fail(VG_E_INVALID, "null edits");
    if (n > (1u << 20)) return fail(VG_E_INVALID, "%u edits (at most 2^20 per call)", n);
    for (int k = 0; k < 3; ++k)
        if (w->dw.origin[k] % 64 != 0)
            return fail(VG_E_UNSUPPORTED, "edits need a chunk-aligned window origin");
    vg_edit_state *s;
    st = edit_state(w, &s);
    if (st) return st;
    hipStream_t hs = (hipStream_t)stream;
    const Dims &d = w->d;
    // validate, and bound the (brick, edit) pairs by the clipped brick boxes
    uint64_t bound = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const vg_edit &e = edits[i];
        if (e.radius > kMaxEditRadius)
            return fail(VG_E_INVALID, "edit %u: radius %u above %d", i, e.radius, kMaxEditRadius);
        for (int k = 0; k < 3; ++k)
            if (e.center[k] < -(1 << 30) || e.center[k] > (1 << 30))
                return fail(VG_E_INVALID, "edit %u: centre out of range", i);
        EditRange r;
        if (!edit_range(e, w->dw.origin, w->dw.dims, r)) continue;
        bound += (uint64_t)(r.hi[0] - r.lo[0] + 1) * (uint64_t)(r.hi[1] - r.lo[1] + 1) *
                 (uint64_t)(r.hi[2] - r.lo[2] + 1);
    }
    if (bound > (1ull << 30))
        return fail(VG_E_INVALID, "edit batch spans %llu brick slots (at most 2^30): split it",
                    (unsigned long long)bound);
    st = ensure_edit_capacity(s, n, (size_t)bound);
    if (st) return st;
    const uint32_t tw = (uint32_t)((d.n_tiles + 31) / 32), cw = (uint32_t)((d.n_chunks + 31) / 32),
                   lw = (uint32_t)((d.n_l4 + 31) / 32);
    const uint32_t mw = std::max(tw, std::max(cw, lw));
    hipLaunchKernelGGL(k_edit_clear, dim3((mw + 255) / 256), dim3(256), 0, hs, s->tile_dirty, tw,
                       s->chunk_dirty, cw, s->l4_dirty, lw, s->tile_n);
    HIP_TRY(hipGetLastError());
    if (bound == 0) return VG_OK;
    // staging[slot] is rewritten only once the copy that last read it has finished
    const int slot = s->slot;
    s->slot ^= 1;
    if (s->staged_rec[slot]) HIP_TRY(hipEventSynchronize(s->staged[slot]));
    std::memcpy(s->staging[slot], edits, (size_t)n * sizeof(vg_edit));
    HIP_TRY(hipMemcpyAsync(s->edits, s->staging[slot], (size_t)n * sizeof(vg_edit),
                           hipMemcpyHostToDevice, hs));
    HIP_TRY(hipEventRecord(s->staged[slot], hs));
    s->staged_rec[slot] = true;
    const dim3 eg((n + 3) / 4);
    hipLaunchKernelGGL(k_edit_enum<false>, eg, dim3(256), 0, hs, w->dw, s->edits, n, s->cnt,
                       s->end, s->list, s->tile_flag, s->tile_list, s->tile_n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(scan_u32(s->cnt, s->end, (uint32_t)d.n_bricks, static_cast<uint32_t *>(s->temp), true,
                     hs));
    hipLaunchKernelGGL(k_edit_enum<true>, eg, dim3(256), 0, hs, w->dw, s->edits, n, s->cnt, s->end,
                       s->list, s->tile_flag, s->tile_list, s->tile_n);
    HIP_TRY(hipGetLastError());
    // touched tiles are at most n_tiles; the list walkers use machine-sized grids
    const uint32_t tb_grid = (uint32_t)std::min<int64_t>(d.n_tiles, 2048);
    hipLaunchKernelGGL(k_edit_bricks, dim3(tb_grid), dim3(1024), 0, hs, w->dw, s->edits, s->end,
                       s->list, s->tile_list, s->tile_n, s->tile_flag, w->occ, w->mat,
                       w->brick_uni, s->tile_dirty);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_edit_tiles, dim3((uint32_t)std::min<int64_t>((d.n_tiles + 3) / 4, 1024)),
                       dim3(256), 0, hs, s->tile_dirty, s->tile_list, s->tile_n, w->occ,
                       w->brick_uni, w->tile, w->tile_uni, s->chunk_dirty, (uint32_t)d.c[0],
                       (uint32_t)d.c[1], w->ytop_dev);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_edit_chunks, dim3((uint32_t)((d.n_chunks + 3) / 4)), dim3(256), 0, hs,
                       s->chunk_dirty, w->dw, (uint32_t)d.n_chunks, w->tile, w->chunk, s->l4_dirty);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_edit_l4, dim3((uint32_t)((d.n_l4 + 255) / 256)), dim3(256), 0, hs,
                       s->l4_dirty, w->dw, (uint32_t)d.n_l4, w->chunk, w->l4);
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

int vg_frame_dirty_blocks(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                          const float sun[3], float shadow_tmax, uint32_t *mask, void *stream) {
    int st = check_world(w);
    if (st) return st;
    if (!camera || !mask) return fail(VG_E_INVALID, "null camera/mask");
    if (width <= 0 || height <= 0) return fail(VG_E_INVALID, "size %dx%d", width, height);
    vg_edit_state *s;
    st = edit_state(w, &s);
    if (st) return st;
    hipStream_t hs = (hipStream_t)stream;
    const uint32_t nbx = (uint32_t)(width + 7) / 8, nby = (uint32_t)(height + 7) / 8;
    const uint32_t nwords = (nbx * nby + 31) / 32;
    hipLaunchKernelGGL(k_mask_clear, dim3((nwords + 1 + 255) / 256), dim3(256), 0, hs, mask,
                       nwords + 1);
    HIP_TRY(hipGetLastError());
    float sd[3] = {0.0f, 0.0f, 0.0f};
    const int shadow = sun != nullptr && host_normalize(sun, sd);
    hipLaunchKernelGGL(k_edit_blocks, dim3((uint32_t)std::min<int64_t>((w->d.n_tiles + 3) / 4, 1024)),
                       dim3(256), 0, hs, w->dw, s->tile_dirty, s->tile_list, s->tile_n,
                       proj_cam(camera, width, height),
                       width, height, nbx, nby, shadow, sd[0], sd[1], sd[2], shadow_tmax, mask,
                       nwords);
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

int vg_frame_mask_words(int32_t width, int32_t height, uint32_t *words_out) {
    if (!words_out || width <= 0 || height <= 0) return fail(VG_E_INVALID, "bad arguments");
    const uint32_t nbx = (uint32_t)(width + 7) / 8, nby = (uint32_t)(height + 7) / 8;
    *words_out = (nbx * nby + 31) / 32 + 1;
    return VG_OK;
}

int vg_world_download(vg_world *w, vg_brick *bricks, uint64_t *tile_mask, uint64_t *chunk_mask,
                      uint64_t *l4_mask, uint8_t *brick_uniform, uint8_t *tile_uniform) {
    int st = check_world(w);
    if (st) return st;
    const Dims &d = w->d;
    HIP_TRY(hipDeviceSynchronize());
    if (l4_mask) HIP_TRY(hipMemcpy(l4_mask, w->l4, (size_t)d.n_l4 * 8, hipMemcpyDeviceToHost));
    if (chunk_mask)
        HIP_TRY(hipMemcpy(chunk_mask, w->chunk, (size_t)d.n_chunks * 8, hipMemcpyDeviceToHost));
    if (tile_mask || tile_uniform) {
        void *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, (size_t)d.n_tiles * 9));
        uint64_t *tref = (uint64_t *)tmp;
        uint8_t *turef = (uint8_t *)tmp + (size_t)d.n_tiles * 8;
        hipLaunchKernelGGL(k_download_tiles, dim3((uint32_t)((d.n_tiles + 255) / 256)), dim3(256), 0,
                           0, w->tile, w->tile_uni, d.n_tiles, d.t[0], d.t[1], d.c[0], d.c[1], tref,
                           turef);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && tile_mask)
            e = hipMemcpy(tile_mask, tref, (size_t)d.n_tiles * 8, hipMemcpyDeviceToHost);
        if (e == hipSuccess && tile_uniform)
            e = hipMemcpy(tile_uniform, turef, (size_t)d.n_tiles, hipMemcpyDeviceToHost);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(VG_E_DEVICE, "tile download: %s", hipGetErrorString(e));
    }
    if (bricks || brick_uniform) {
        const int64_t slab = std::min<int64_t>(d.n_bricks, (int64_t)1 << 22);
        void *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, (size_t)slab * (sizeof(vg_brick) + 1)));
        vg_brick *sb = (vg_brick *)tmp;
        uint8_t *sbu = (uint8_t *)tmp + (size_t)slab * sizeof(vg_brick);
        hipError_t e = hipSuccess;
        for (int64_t first = 0; first < d.n_bricks && e == hipSuccess; first += slab) {
            const int64_t cnt = std::min(slab, d.n_bricks - first);
            hipLaunchKernelGGL(k_download_bricks, dim3((uint32_t)((cnt + 255) / 256)), dim3(256), 0, 0,
                               w->occ, w->mat, w->brick_uni, first, cnt, d.b[0], d.b[1], d.c[0],
                               d.c[1], sb, sbu);
            e = hipGetLastError();
            if (e == hipSuccess && bricks)
                e = hipMemcpy(bricks + first, sb, (size_t)cnt * sizeof(vg_brick), hipMemcpyDeviceToHost);
            if (e == hipSuccess && brick_uniform)
                e = hipMemcpy(brick_uniform + first, sbu, (size_t)cnt, hipMemcpyDeviceToHost);
        }
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(VG_E_DEVICE, "brick download: %s", hipGetErrorString(e));
    }
    return VG_OK;
}

int vg_world_generate(vg_world *w, uint64_t seed, const int32_t origin[3], void *stream) {
    if (!w || !origin) return fail(VG_E_INVALID, "null world/origin");
    const Dims &d = w->d;
    for (int k = 0; k < 3; ++k)
        if (pos_mod(origin[k], d.v[k]) != 0 && d.v[k] % 256 != 0)
            return fail(VG_E_INVALID,
                        "origin[%d]=%d folds the window but dims[%d]=%d is not a multiple of 256", k,
                        origin[k], k, d.v[k]);
    int st = set_device(w);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize()); // kernels on any stream may still read the old world
    hipStream_t hs = (hipStream_t)stream;
    w->uploaded = false;
    w->dw.fold = 0;
    for (int k = 0; k < 3; ++k) {
        w->dw.origin[k] = origin[k];
        w->dw.ro_mod[k] = pos_mod(origin[k], d.v[k]);
        if (w->dw.ro_mod[k]) w->dw.fold = 1;
    }
    // seed_offset_xz (src/voxel.rs:1163-1168), as the host generator
    const float sx = (float)(int32_t)(uint32_t)(seed >> 32) * 0.01734f;
    const float sz = (float)(int32_t)(uint32_t)(seed & 0xFFFFFFFFu) * 0.02153f;
    const uint32_t ncols = (uint32_t)((d.v[0] >> 2) * (d.v[2] >> 2));
    hipLaunchKernelGGL(k_gen_bricks, dim3((ncols + 3) / 4), dim3(256), 0, hs, w->dw, sx, sz, ncols,
                       w->occ, w->mat, w->brick_uni);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_set_ytop, dim3(1), dim3(1), 0, hs, w->ytop_dev,
                       w->dw.ro_mod[1] ? d.v[1] : 0);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_gen_tiles, dim3((uint32_t)((d.n_tiles + 3) / 4)), dim3(256), 0, hs,
                       (uint32_t)d.n_tiles, w->occ, w->brick_uni, w->tile, w->tile_uni,
                       (uint32_t)d.c[0], (uint32_t)d.c[1], w->ytop_dev);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_gen_chunks, dim3((uint32_t)((d.n_chunks + 3) / 4)), dim3(256), 0, hs,
                       (uint32_t)d.n_chunks, w->tile, w->chunk);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_gen_l4, dim3((uint32_t)((d.n_l4 + 255) / 256)), dim3(256), 0, hs, w->dw,
                       (uint32_t)d.n_l4, w->chunk, w->l4);
    HIP_TRY(hipGetLastError());
    w->uploaded = true;
    return VG_OK;
}

} // extern "C"

namespace vg {

// Dirty-block list of a mask for the masked frame (vg_device.hip): list + count on device.
int edit_block_list(vg_world *w, const uint32_t *mask, uint32_t nblocks, hipStream_t stream,
                    const uint32_t **list, const uint32_t **list_n) {
    vg_edit_state *s;
    int st = edit_state(w, &s);
    if (st) return st;
    if (nblocks > s->blist_cap) {
        HIP_TRY(hipDeviceSynchronize());
        if (s->blist) (void)hipFree(s->blist);
        s->blist = nullptr;
        s->blist_cap = 0;
        if (hipMalloc((void **)&s->blist, (size_t)nblocks * 4) != hipSuccess)
            return fail(VG_E_NOMEM, "block list");
        s->blist_cap = nblocks;
    }
    const uint32_t nwords = (nblocks + 31) / 32;
    hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(1), 0, stream, s->blist_n);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_mask_compact, dim3((nblocks + 255) / 256), dim3(256), 0, stream, mask, nwords,
                       nblocks, s->blist, s->blist_n);
    HIP_TRY(hipGetLastError());
    *list = s->blist;
    *list_n = s->blist_n;
    return VG_OK;
}

} // namespace vg
