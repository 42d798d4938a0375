// Dynamic world on gfx950: sphere edits of an uploaded world, applied on the device with the
// reference's sequential semantics (apply_sphere -> apply_edit -> set_voxel, src/app.rs:199-210,
// src/voxel.rs:450-474, 871-946), an incremental bottom-up rebuild of the pyramid and uniform
// tables, and the 8x8 screen blocks a frame must retrace afterwards (the reference's
// project_brick_to_tiles, src/temporal.rs:16-111, made conservative).
//
//   pairs   the host sizes, per edit, the bricks its sphere touches inside the window; a
//           kernel emits (nested brick index, edit index) pairs edit by edit; a stable
//           radix sort (rocPRIM through hipCUB) groups them by brick keeping edit order.
//   bricks  one wave per touched brick, one lane per voxel: the lane applies the brick's
//           spheres in edit order (the last covering sphere wins, exactly as sequential
//           set_voxel calls), the occupancy word is the wave's ballot, the brick-uniform byte
//           a wave-wide equality; changed bricks mark their tile dirty.
//   tiles / chunks / L4   one wave per dirty node: the node word is the ballot of its 64
//           children being non-empty, the tile-uniform byte a wave-wide equality.
//   blocks  one thread per changed tile projects the tile's box (primary rays) and the box
//           swept against the sun over the shadow range (shadow rays) onto the screen and
//           marks the covered 8x8 blocks (conservative: 2-pixel margins, the whole screen
//           when a box straddles the camera plane).
#include "vg_world.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

struct vg_edit_state {
    size_t cap_edits, cap_pairs;
    vg_edit *edits;          // device copy of the last batch
    uint32_t *off;           // first pair of each edit (n + 1)
    uint32_t *pk_in, *pk_out; // pair keys: nested brick index
    uint32_t *pv_in, *pv_out; // pair values: edit index
    void *temp;
    size_t temp_bytes;
    uint32_t *tile_dirty;    // bit per nested tile changed by the last batch
    uint32_t *chunk_dirty;   // bit per chunk
    uint32_t *l4_dirty;      // bit per L4 cell
    uint32_t *blist;         // dirty screen blocks of a masked frame
    uint32_t *blist_n;
    size_t blist_cap;
    vg_edit *staging;        // pinned host staging of the edit batch
    uint32_t *staging_off;
    size_t staging_cap;
    int key_bits;
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
                             uint32_t nc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0u;
    if (i < nb) b[i] = 0u;
    if (i < nc) c[i] = 0u;
}

__global__ void k_edit_pairs(DevWorld w, const vg_edit *__restrict__ edits, uint32_t n,
                             const uint32_t *__restrict__ off, uint32_t *__restrict__ keys,
                             uint32_t *__restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const vg_edit e = edits[i];
    EditRange r;
    if (!edit_range(e, w.origin, w.dims, r)) return;
    uint32_t o = off[i];
    const uint32_t end = off[i + 1];
    for (int32_t y = r.lo[1]; y <= r.hi[1]; ++y)
        for (int32_t z = r.lo[2]; z <= r.hi[2]; ++z)
            for (int32_t x = r.lo[0]; x <= r.hi[0]; ++x) {
                const int32_t wb[3] = {x, y, z};
                if (!sphere_touches_brick(e, wb)) continue;
                int32_t sb[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    int32_t s = (wb[k] * 4) % w.dims[k];
                    s = s < 0 ? s + w.dims[k] : s;
                    sb[k] = s >> 2;
                }
                if (o < end) { // the host sized exactly this many
                    keys[o] = brick_nested(w, sb);
                    vals[o] = i;
                }
                ++o;
            }
}

// One wave per chunk of 64 sorted pairs: every brick whose first pair lies in the chunk is
// processed by this wave, its spheres in edit order.
__global__ __launch_bounds__(256) void k_edit_bricks(DevWorld w, const vg_edit *__restrict__ edits,
                                                     const uint32_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ vals,
                                                     uint32_t n_pairs, uint64_t *__restrict__ occ,
                                                     uint8_t *__restrict__ mat,
                                                     uint8_t *__restrict__ bu,
                                                     uint32_t *__restrict__ tile_dirty) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t chunk = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t p = chunk * 64u + lane;
    if (chunk * 64u >= n_pairs) return; // wave-uniform
    const bool head = p < n_pairs && (p == 0 || keys[p] != keys[p - 1]);
    uint64_t heads = __ballot(head);
    // voxel of this lane inside a brick: bit x + 4z + 16y
    const int32_t vx = (int32_t)(lane & 3u), vy = (int32_t)(lane >> 4), vz = (int32_t)((lane >> 2) & 3u);
    while (heads) {
        const int src = __ffsll((long long)heads) - 1;
        heads &= heads - 1;
        const uint32_t hp = chunk * 64u + (uint32_t)src;
        const uint32_t bn = keys[hp];
        int32_t sb[3];
        brick_slot(w, bn, sb);
        const int32_t wv[3] = {slot_to_world(sb[0] * 4 + vx, w.origin[0], w.dims[0]),
                               slot_to_world(sb[1] * 4 + vy, w.origin[1], w.dims[1]),
                               slot_to_world(sb[2] * 4 + vz, w.origin[2], w.dims[2])};
        const uint64_t old_occ = occ[bn];
        const bool old_bit = (old_occ >> lane) & 1u;
        const uint32_t raw = mat[(size_t)bn * 64u + lane];
        // a voxel some sphere covers takes the last covering sphere's material with
        // Brick::set semantics (src/voxel.rs:315-325); the others keep their bit and byte
        bool touched = false;
        uint32_t m = raw;
        for (uint32_t q = hp; q < n_pairs && keys[q] == bn; ++q) {
            const vg_edit e = edits[vals[q]];
            const int32_t dx = wv[0] - e.center[0], dy = wv[1] - e.center[1], dz = wv[2] - e.center[2];
            if (dx * dx + dy * dy + dz * dz <= (int32_t)e.radius * (int32_t)e.radius) {
                m = e.material;
                touched = true;
            }
        }
        const bool bit = touched ? m != 0u : old_bit;
        const uint64_t new_occ = __ballot(bit);
        const bool changed = __ballot(touched && (m != raw || bit != old_bit)) != 0;
        const uint32_t m0 = __shfl(m, 0, 64);
        const bool all_same = __ballot(m == m0) == ~0ull;
        if (touched) mat[(size_t)bn * 64u + lane] = (uint8_t)m;
        if (lane == 0) {
            occ[bn] = new_occ;
            bu[bn] = (new_occ == ~0ull && all_same && m0 != 0u) ? (uint8_t)m0 : (uint8_t)0;
            if (changed) atomicOr(tile_dirty + ((bn >> 6) >> 5), 1u << ((bn >> 6) & 31u));
        }
    }
}

// One wave per nested tile: a dirty tile recomputes its word and uniform byte from its 64
// bricks and marks its chunk.
__global__ __launch_bounds__(256) void k_edit_tiles(const uint32_t *__restrict__ tile_dirty,
                                                    uint32_t n_tiles, const uint64_t *__restrict__ occ,
                                                    const uint8_t *__restrict__ bu,
                                                    uint64_t *__restrict__ tile,
                                                    uint8_t *__restrict__ tu,
                                                    uint32_t *__restrict__ chunk_dirty) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tn = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (tn >= n_tiles || !((tile_dirty[tn >> 5] >> (tn & 31u)) & 1u)) return; // wave-uniform
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

__device__ __forceinline__ void mark_rect(uint32_t *mask, uint32_t nbx, uint32_t nby, int32_t width,
                                          int32_t height, const float rect[4]) {
    const float xa = fmaxf(rect[0], 0.0f), ya = fmaxf(rect[1], 0.0f);
    const float xb = fminf(rect[2], (float)(width - 1)), yb = fminf(rect[3], (float)(height - 1));
    if (xa > xb || ya > yb) return;
    const uint32_t bx0 = (uint32_t)xa / 8u, bx1 = (uint32_t)xb / 8u;
    const uint32_t by0 = (uint32_t)ya / 8u, by1 = (uint32_t)yb / 8u;
    for (uint32_t by = by0; by <= by1 && by < nby; ++by) {
        uint32_t b = by * nbx + bx0;
        const uint32_t end = by * nbx + (bx1 < nbx ? bx1 : nbx - 1u);
        while (b <= end) {
            const uint32_t word = b >> 5, first = b & 31u;
            const uint32_t last = (end >> 5) == word ? (end & 31u) : 31u;
            const uint32_t bits = (last == 31u ? ~0u : ((1u << (last + 1u)) - 1u)) & ~((1u << first) - 1u);
            atomicOr(mask + word, bits);
            b = word * 32u + last + 1u;
        }
    }
}

// One thread per nested tile changed by the last edit batch: mark the 8x8 blocks whose
// primary rays can enter the tile box, and (sun set) whose primary hits can cast a shadow ray
// through it. mask[nwords] is the whole-screen flag.
__global__ void k_edit_blocks(DevWorld w, const uint32_t *__restrict__ tile_dirty, uint32_t n_tiles,
                              ProjCam c, int32_t width, int32_t height, uint32_t nbx, uint32_t nby,
                              int shadow, float sx, float sy, float sz, float stmax,
                              uint32_t *__restrict__ mask, uint32_t nwords) {
    const uint32_t tn = blockIdx.x * blockDim.x + threadIdx.x;
    if (tn >= n_tiles || !((tile_dirty[tn >> 5] >> (tn & 31u)) & 1u)) return;
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
        mask[nwords] = 1u;
        return;
    }
    if (r == 1) mark_rect(mask, nbx, nby, width, height, rect);
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
        mask[nwords] = 1u;
        return;
    }
    if (r == 1) mark_rect(mask, nbx, nby, width, height, rect);
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

// --------------------------------------------------------------------- host --
static void free_edit_state(vg_edit_state *s) {
    void *dev[] = {s->edits, s->off, s->pk_in, s->pk_out, s->pv_in, s->pv_out, s->temp,
                   s->tile_dirty, s->chunk_dirty, s->l4_dirty, s->blist, s->blist_n};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (s->staging) (void)hipHostFree(s->staging);
    if (s->staging_off) (void)hipHostFree(s->staging_off);
    delete s;
}

static int edit_state(vg_world *w, vg_edit_state **out) {
    if (!w->edit) {
        vg_edit_state *s = new vg_edit_state();
        std::memset(s, 0, sizeof(*s));
        const Dims &d = w->d;
        s->key_bits = 1;
        while ((1ll << s->key_bits) < d.n_bricks) ++s->key_bits;
        const size_t tw = (size_t)(d.n_tiles + 31) / 32, cw = (size_t)(d.n_chunks + 31) / 32,
                     lw = (size_t)(d.n_l4 + 31) / 32;
        bool ok = hipMalloc((void **)&s->tile_dirty, tw * 4) == hipSuccess &&
                  hipMalloc((void **)&s->chunk_dirty, cw * 4) == hipSuccess &&
                  hipMalloc((void **)&s->l4_dirty, lw * 4) == hipSuccess &&
                  hipMalloc((void **)&s->blist_n, 256) == hipSuccess;
        if (ok) ok = hipMemset(s->tile_dirty, 0, tw * 4) == hipSuccess &&
                     hipMemset(s->chunk_dirty, 0, cw * 4) == hipSuccess &&
                     hipMemset(s->l4_dirty, 0, lw * 4) == hipSuccess;
        if (!ok) {
            free_edit_state(s);
            return fail(VG_E_NOMEM, "edit state");
        }
        w->edit = s;
    }
    *out = w->edit;
    return VG_OK;
}

static int ensure_edit_capacity(vg_edit_state *s, size_t n_edits, size_t n_pairs) {
    if (n_edits > s->cap_edits || n_edits + 1 > s->staging_cap) {
        HIP_TRY(hipDeviceSynchronize());
        if (s->edits) (void)hipFree(s->edits);
        if (s->off) (void)hipFree(s->off);
        if (s->staging) (void)hipHostFree(s->staging);
        if (s->staging_off) (void)hipHostFree(s->staging_off);
        s->edits = nullptr;
        s->off = nullptr;
        s->staging = nullptr;
        s->staging_off = nullptr;
        s->cap_edits = s->staging_cap = 0;
        const size_t cap = std::max<size_t>(n_edits, 4096);
        if (hipMalloc((void **)&s->edits, cap * sizeof(vg_edit)) != hipSuccess ||
            hipMalloc((void **)&s->off, (cap + 1) * sizeof(uint32_t)) != hipSuccess ||
            hipHostMalloc((void **)&s->staging, cap * sizeof(vg_edit)) != hipSuccess ||
            hipHostMalloc((void **)&s->staging_off, (cap + 1) * sizeof(uint32_t)) != hipSuccess)
            return fail(VG_E_NOMEM, "edit buffers for %zu edits", cap);
        s->cap_edits = cap;
        s->staging_cap = cap + 1;
    }
    if (n_pairs > s->cap_pairs) {
        HIP_TRY(hipDeviceSynchronize());
        void *ptrs[] = {s->pk_in, s->pk_out, s->pv_in, s->pv_out, s->temp};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        s->pk_in = s->pk_out = s->pv_in = s->pv_out = nullptr;
        s->temp = nullptr;
        s->cap_pairs = 0;
        const size_t cap = std::max<size_t>(n_pairs + n_pairs / 4, 1u << 16);
        bool ok = hipMalloc((void **)&s->pk_in, cap * 4) == hipSuccess &&
                  hipMalloc((void **)&s->pk_out, cap * 4) == hipSuccess &&
                  hipMalloc((void **)&s->pv_in, cap * 4) == hipSuccess &&
                  hipMalloc((void **)&s->pv_out, cap * 4) == hipSuccess;
        if (!ok) return fail(VG_E_NOMEM, "edit pairs (%zu)", cap);
        size_t tb = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, s->pk_in, s->pk_out, s->pv_in,
                                                   s->pv_out, (int)cap, 0, s->key_bits));
        if (hipMalloc(&s->temp, tb) != hipSuccess) return fail(VG_E_NOMEM, "sort scratch");
        s->temp_bytes = tb;
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
    // pair counts on the host (exact), validated edits
    std::vector<uint32_t> off((size_t)n + 1, 0);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (edits[i].radius > kMaxEditRadius)
            return fail(VG_E_INVALID, "edit %u: radius %u above %d", i, edits[i].radius,
                        kMaxEditRadius);
        off[i] = (uint32_t)total;
        EditRange r;
        if (!edit_range(edits[i], w->dw.origin, w->dw.dims, r)) continue;
        for (int32_t y = r.lo[1]; y <= r.hi[1]; ++y)
            for (int32_t z = r.lo[2]; z <= r.hi[2]; ++z)
                for (int32_t x = r.lo[0]; x <= r.hi[0]; ++x) {
                    const int32_t wb[3] = {x, y, z};
                    if (sphere_touches_brick(edits[i], wb)) ++total;
                }
        if (total > 0x7FFFFFFFull) return fail(VG_E_INVALID, "too many touched bricks");
    }
    off[n] = (uint32_t)total;
    st = ensure_edit_capacity(s, n, (size_t)total);
    if (st) return st;
    const uint32_t tw = (uint32_t)((d.n_tiles + 31) / 32), cw = (uint32_t)((d.n_chunks + 31) / 32),
                   lw = (uint32_t)((d.n_l4 + 31) / 32);
    const uint32_t mw = std::max(tw, std::max(cw, lw));
    hipLaunchKernelGGL(k_edit_clear, dim3((mw + 255) / 256), dim3(256), 0, hs, s->tile_dirty, tw,
                       s->chunk_dirty, cw, s->l4_dirty, lw);
    HIP_TRY(hipGetLastError());
    if (total == 0) return VG_OK;
    // the staging buffers are reused: the previous batch's copies must have finished
    HIP_TRY(hipStreamSynchronize(hs));
    std::memcpy(s->staging, edits, (size_t)n * sizeof(vg_edit));
    std::memcpy(s->staging_off, off.data(), off.size() * sizeof(uint32_t));
    HIP_TRY(hipMemcpyAsync(s->edits, s->staging, (size_t)n * sizeof(vg_edit),
                           hipMemcpyHostToDevice, hs));
    HIP_TRY(hipMemcpyAsync(s->off, s->staging_off, off.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, hs));
    hipLaunchKernelGGL(k_edit_pairs, dim3((n + 255) / 256), dim3(256), 0, hs, w->dw, s->edits, n,
                       s->off, s->pk_in, s->pv_in);
    HIP_TRY(hipGetLastError());
    size_t tb = s->temp_bytes;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(s->temp, tb, s->pk_in, s->pk_out, s->pv_in,
                                               s->pv_out, (int)total, 0, s->key_bits, hs));
    const uint32_t chunks = (uint32_t)((total + 63) / 64);
    hipLaunchKernelGGL(k_edit_bricks, dim3((chunks + 3) / 4), dim3(256), 0, hs, w->dw, s->edits,
                       s->pk_out, s->pv_out, (uint32_t)total, w->occ, w->mat, w->brick_uni,
                       s->tile_dirty);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_edit_tiles, dim3((uint32_t)((d.n_tiles + 3) / 4)), dim3(256), 0, hs,
                       s->tile_dirty, (uint32_t)d.n_tiles, w->occ, w->brick_uni, w->tile,
                       w->tile_uni, s->chunk_dirty);
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
    hipLaunchKernelGGL(k_edit_blocks, dim3((uint32_t)((w->d.n_tiles + 255) / 256)), dim3(256), 0, hs,
                       w->dw, s->tile_dirty, (uint32_t)w->d.n_tiles, proj_cam(camera, width, height),
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
