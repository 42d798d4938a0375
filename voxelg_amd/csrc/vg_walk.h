// Device-side world view and the exact hierarchical voxel walk shared by every kernel
// file of the engine (trace, probes, edits).
//
// HBM layout (all arrays owned by vg_world):
//   l4[n_l4]          u64  reference order (x fastest)             one bit per child chunk
//   chunk[n_chunks]   u64  reference order                         one bit per child tile
//   tile[c*64 + tb]   u64  nested: chunk index, then tile bit      one bit per child brick
//   tile_uni[same]    u8
//   occ[t*64 + bb]    u64  nested: nested tile, then brick bit     one bit per voxel
//   brick_uni[same]   u8
//   mat[b*64 + vb]    u8   nested brick, then voxel bit
// so the 64 child words of any tile (or chunk) are one contiguous 512-byte run that a
// wave reads in one coalesced pass, and a brick's 64 materials are one 64-byte line.
//
// Exactness: the walk uses only IEEE +, -, *, / and floor in fp32 (contraction off,
// correctly rounded divide and sqrt), recomputes every plane-crossing t from integer
// coordinates, and skips empty cells with an integer-exact exit computation, so its
// result equals a one-voxel-at-a-time walk bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "vg_internal.h"

namespace vg {

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(VG_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

struct DevWorld {
    const uint64_t *l4;
    const uint64_t *chunk;
    const uint64_t *tile;
    const uint64_t *occ;
    const uint8_t *tile_uni;
    const uint8_t *brick_uni;
    const uint8_t *mat;
    int32_t dims[3];
    int32_t ro_mod[3];
    int32_t origin[3];
    int32_t cdim[3];
    int32_t l4dim[3];
    int32_t fold; // 1 when the window origin folds any axis (toroidal slots)
    // Every occupied voxel has a window-local y below ytop, so a ray at y >= ytop that does
    // not move down can hit nothing: the walk ends there with a miss. Kernels load the
    // world's current bound from ytop_p (kept by upload, generation and edits) with
    // world_ready; the host-side value is dims[1], which never ends a walk early.
    int32_t ytop;
    const int32_t *ytop_p;
};

// The world's y bound for this launch (wave-uniform: one scalar load per kernel).
__device__ __forceinline__ void world_ready(DevWorld &w) { w.ytop = *w.ytop_p; }

struct CamParams {
    float origin[3];
    float forward[3];
    float right[3];
    float up[3];
    float tan_half_fov;
    float res[2];
    float jitter[2];
};

// ------------------------------------------------------------ material rules --
// shaders/raymarch.wgsl:891-920 predicates; src/raycast.rs:88-89 fringe exclusion.
__device__ __forceinline__ bool is_transparent(uint32_t m) { return (m >= 5u && m <= 12u) || m == 18u; }
__device__ __forceinline__ bool is_foliage(uint32_t m) {
    return m == 14u || m == 25u || m == 26u || m == 27u || m == 30u || m == 31u || m == 33u ||
           m == 34u || m == 35u || m == 36u || m == 37u;
}
__device__ __forceinline__ bool uniform_optimisable(uint32_t m) {
    return !is_foliage(m) && !is_transparent(m);
}
__device__ __forceinline__ bool solid_mat(uint32_t m, uint32_t flags) {
    if (m == 0u || m == 33u) return false;
    if ((flags & VG_RAY_SKIP_TRANSPARENT) && is_transparent(m)) return false;
    if ((flags & VG_RAY_ANY_HIT) && m == 35u) return false;
    return true;
}

__device__ __forceinline__ float fsel_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fsel_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float safe_inv(float x) { return fabsf(x) < 1e-8f ? 1e30f : 1.0f / x; }
__device__ __forceinline__ int cbit(int32_t x, int32_t y, int32_t z) { return x + z * 4 + y * 16; }

// --------------------------------------------------------------- walk state --
struct Walk {
    float org[3], dir[3], inv[3];
    float tn[3];     // t of the next plane crossing on each axis (+inf for an axis that never steps)
    int32_t step[3];
    int32_t l[3];    // current voxel, window-local
    float t_cur;     // entry t of the current voxel
    int32_t last_axis;
    int32_t entry_face;
    int32_t face_pack; // 4 bits per axis: face code of an entry through that axis
};

struct HitRec {
    int32_t v[3];
    float t;
    uint32_t face, mat, hit;
};

__device__ __forceinline__ void hit_miss(HitRec &h) {
    h.v[0] = h.v[1] = h.v[2] = 0;
    h.t = -1.0f;
    h.face = VG_FACE_NONE;
    h.mat = 0;
    h.hit = 0;
}

// t of the plane crossed to enter voxel e along axis k (plane e when stepping up, e + 1
// when stepping down), always from the integer coordinate.
__device__ __forceinline__ float t_plane(const Walk &r, int k, int32_t e) {
    int32_t p = r.step[k] > 0 ? e : e + 1;
    return ((float)p - r.org[k]) * r.inv[k];
}

__device__ __forceinline__ float t_next(const Walk &r, int k) {
    return r.step[k] != 0 ? t_plane(r, k, r.l[k] + r.step[k]) : INFINITY;
}

// Ray setup: window-local origin, slab entry, 1e-3 bias clamped into the window
// (shaders/raymarch.wgsl:3369-3411, src/raycast.rs:34-50). dir is already normalized.
// `inv` holds safe_inv of each component of d (a caller tracing many rays along one
// direction computes it once).
__device__ __forceinline__ bool walk_init(Walk &r, const DevWorld &w, const float o[3],
                                          const float d[3], const float inv[3]) {
    float tmn[3], tmx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.org[k] = o[k] - (float)w.origin[k];
        r.dir[k] = d[k];
        r.inv[k] = inv[k];
        float t0 = (0.0f - r.org[k]) * r.inv[k];
        float t1 = ((float)w.dims[k] - r.org[k]) * r.inv[k];
        tmn[k] = fsel_min(t0, t1);
        tmx[k] = fsel_max(t0, t1);
        r.step[k] = fabsf(d[k]) < 1e-8f ? 0 : (d[k] > 0.0f ? 1 : -1);
    }
    if (r.step[0] == 0 && r.step[1] == 0 && r.step[2] == 0) return false; // cannot advance
    float t_enter = fsel_max(fsel_max(tmn[0], fsel_max(tmn[1], tmn[2])), 0.0f);
    float t_exit = fsel_min(tmx[0], fsel_min(tmx[1], tmx[2]));
    if (t_enter >= t_exit || t_exit < 0.0f) return false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float p = r.org[k] + d[k] * (t_enter + 1e-3f);
        p = fsel_min(fsel_max(p, 1e-3f), (float)w.dims[k] - 1e-3f);
        r.l[k] = (int32_t)floorf(p);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) r.tn[k] = t_next(r, k);
    r.t_cur = t_enter;
    r.last_axis = -1;
    r.face_pack = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int f = r.step[k] > 0 ? 2 * k + 1 : (r.step[k] < 0 ? 2 * k : VG_FACE_NONE);
        r.face_pack |= f << (4 * k);
    }
    int la = (tmn[0] >= tmn[1] && tmn[0] >= tmn[2]) ? 0 : (tmn[1] >= tmn[2] ? 1 : 2);
    r.entry_face = (r.face_pack >> (4 * la)) & 15;
    return true;
}

__device__ __forceinline__ bool walk_init(Walk &r, const DevWorld &w, const float o[3],
                                          const float d[3]) {
    const float inv[3] = {safe_inv(d[0]), safe_inv(d[1]), safe_inv(d[2])};
    return walk_init(r, w, o, d, inv);
}

// Advance out of the aligned cell of size S (1 = just the current voxel) that holds the
// current voxel, to exactly the state a one-voxel-at-a-time walk reaches at its first
// crossing of the cell's faces (T*, ax*): crossings are processed in (t, axis) order, ties
// to the lowest axis (src/raycast.rs:92-103), every t from its integer plane. The crossed
// axis lands just outside the face; another axis moves only if its cached next crossing
// orders before (T*, ax*) - then it lands on the last in-cell plane that does, found from a
// floor estimate corrected by direct t comparisons (t along an axis is monotone in the
// plane coordinate), and the first rejected plane's t becomes its next crossing.
// For S = 1 this is the plain voxel step; for S > 1 it replaces skip_to_cell
// (shaders/raymarch.wgsl:4807-4864) with an integer-exact exit.
__device__ __forceinline__ void walk_advance(Walk &r, const int32_t sl[3], int32_t S) {
    int32_t lo[3];
    float tf[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = r.l[k] - (sl[k] & (S - 1));
        int32_t plane = r.step[k] > 0 ? lo[k] + S : lo[k];
        tf[k] = r.step[k] != 0 ? ((float)plane - r.org[k]) * r.inv[k] : INFINITY;
    }
    int ax = 0;
    float ts = tf[0];
    if (tf[1] < ts) {
        ts = tf[1];
        ax = 1;
    }
    if (tf[2] < ts) {
        ts = tf[2];
        ax = 2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k == ax) {
            r.l[k] = r.step[k] > 0 ? lo[k] + S : lo[k] - 1;
            r.tn[k] = t_plane(r, k, r.l[k] + r.step[k]);
        } else if (r.tn[k] < ts || (r.tn[k] == ts && k < ax)) {
            // this axis crosses at least one in-cell plane before the exit (so S > 1)
            const int32_t st = r.step[k];
            const int32_t first = r.l[k];
            const int32_t last = st > 0 ? lo[k] + S - 1 : lo[k];
            int32_t e = (int32_t)floorf(r.org[k] + r.dir[k] * ts);
            if (st > 0) {
                e = e < first ? first : e;
                e = e > last ? last : e;
            } else {
                e = e > first ? first : e;
                e = e < last ? last : e;
            }
            // the estimate checked from both sides at once: the crossing into e is before
            // the exit (or e is the first voxel) and the crossing out of it is not (or e is
            // the last in-cell voxel, whose crossing out is the cell face tf[k]); then the
            // corrections below would change nothing and e, with that crossing, is the result
            const float te = t_plane(r, k, e), tx = t_plane(r, k, e + st);
            const bool lo_ok = e == first || te < ts || (te == ts && k < ax);
            const bool hi_ok = e == last || !(tx < ts || (tx == ts && k < ax));
            float nxt = e == last ? tf[k] : tx;
            if (!(lo_ok && hi_ok)) {
                // bounded by the cell width
                for (int g = 0; g < 258 && e != first; ++g) {
                    float t = t_plane(r, k, e);
                    if (t < ts || (t == ts && k < ax)) break;
                    e -= st;
                }
                nxt = tf[k]; // the cell face, if e ends on the last in-cell voxel
                for (int g = 0; g < 258 && e != last; ++g) {
                    float t = t_plane(r, k, e + st);
                    if (!(t < ts || (t == ts && k < ax))) {
                        nxt = t;
                        break;
                    }
                    e += st;
                }
            }
            r.l[k] = e;
            r.tn[k] = nxt;
        }
    }
    r.t_cur = ts;
    r.last_axis = ax;
}

// Jump to the state reached once every plane crossing with t < tb has been processed (a
// prefix of the (t, axis) order): each axis lands on its last plane before tb (floor
// estimate corrected by direct t comparisons), the entry t is the latest crossing, ties to
// the higher axis (processed last). Used after the beam prologue proved every voxel the
// ray visits before tb empty, so the walk continues exactly as if it had stepped there.
__device__ __forceinline__ void walk_advance_to(Walk &r, float tb) {
    float tl = -INFINITY;
    int la = -1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (r.step[k] != 0 && r.tn[k] < tb) {
            const int32_t st = r.step[k];
            const int32_t nearest = r.l[k] + st;
            int32_t e = (int32_t)floorf(r.org[k] + r.dir[k] * tb);
            if (st > 0 ? e < nearest : e > nearest) e = nearest;
            // the estimate checked from both sides at once; the corrections only when it is off
            float tk = t_plane(r, k, e), tx = t_plane(r, k, e + st);
            if (!((e == nearest || tk < tb) && !(tx < tb))) {
                for (int g = 0; g < 4096 && e != nearest && !(t_plane(r, k, e) < tb); ++g)
                    e -= st;
                for (int g = 0; g < 4096 && t_plane(r, k, e + st) < tb; ++g) e += st;
                tk = t_plane(r, k, e);
                tx = t_plane(r, k, e + st);
            }
            r.l[k] = e;
            r.tn[k] = tx;
            if (tk > tl || (tk == tl && k > la)) {
                tl = tk;
                la = k;
            }
        }
    }
    if (la >= 0) {
        r.t_cur = tl;
        r.last_axis = la;
    }
}

__device__ __forceinline__ void emit(const Walk &r, const DevWorld &w, uint32_t m, HitRec &h) {
    h.hit = 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) h.v[k] = r.l[k] + w.origin[k];
    h.t = r.t_cur;
    h.mat = m;
    h.face = r.last_axis < 0 ? (uint32_t)r.entry_face
                             : (uint32_t)((r.face_pack >> (4 * r.last_axis)) & 15);
}

// shaders/raymarch.wgsl:253-295: topmost non-turf voxel's material, layer y = 3 first.
__device__ __forceinline__ uint32_t topmost_material(const DevWorld &w, uint32_t bn, uint64_t occ) {
#pragma unroll
    for (int y = 3; y >= 0; --y) {
        uint32_t layer = (uint32_t)(occ >> (16 * y)) & 0xFFFFu;
        if (layer) {
            uint32_t m = w.mat[(uint64_t)bn * 64u + 16u * y + __builtin_ctz(layer)];
            if (m != 35u) return m;
        }
    }
    return occ ? 35u : 0u;
}

// The 8 child bits of the 2x2x2 group of children holding slot coordinate sl inside a
// parent u64 whose children are 2^(sh-1) voxels wide (bits lx + 4 lz + 16 ly over the
// group's octant). An empty group is an aligned empty cube twice the child size, left
// exactly like any other empty cell; it is not a level of the reference's pyramid.
__device__ __forceinline__ uint64_t group_mask(const int32_t sl[3], int sh) {
    const int ox = (sl[0] >> sh) & 1, oy = (sl[1] >> sh) & 1, oz = (sl[2] >> sh) & 1;
    return (uint64_t)0x330033u << (2 * ox + 8 * oz + 32 * oy);
}

#define LOD_BRICK_T 400.0f
#define TILE_LOD_T 520.0f

// Linear indices of the L4 cell and the chunk holding slot coordinate sl (x fastest, then
// y, then z). Worlds are at most 16384 voxels per axis, so every factor and partial sum stays
// below 2^24 (256 chunks per axis, 2^24 chunks) and full-rate 24-bit multiplies are exact.
__device__ __forceinline__ uint32_t l4_index(const DevWorld &w, const int32_t sl[3]) {
    return (uint32_t)(sl[0] >> 8) + __umul24((uint32_t)(sl[1] >> 8), (uint32_t)w.l4dim[0]) +
           __umul24((uint32_t)(sl[2] >> 8), (uint32_t)(w.l4dim[0] * w.l4dim[1]));
}
__device__ __forceinline__ uint32_t chunk_index(const DevWorld &w, const int32_t sl[3]) {
    return (uint32_t)(sl[0] >> 6) + __umul24((uint32_t)(sl[1] >> 6), (uint32_t)w.cdim[0]) +
           __umul24((uint32_t)(sl[2] >> 6), (uint32_t)(w.cdim[0] * w.cdim[1]));
}

// The walk: phase 1 leaves solid voxels the ray starts in (src/raycast.rs:108-115),
// phase 2 returns the first solid voxel. Each iteration tests the coarsest level and
// skips the largest empty aligned cell (shaders/raymarch.wgsl:4626-4775, plus exact 2x2x2
// groups of empty children). COUNT tallies iteration classes for algorithmic-byte
// accounting.
// The phase a ray carries from one walk iteration to the next besides its Walk.
struct WalkPhase {
    bool phase1;
};

__device__ __forceinline__ void phase_reset(WalkPhase &q) { q.phase1 = true; }

// One iteration of the walk. Returns true when the walk has ended (h then holds the hit,
// or stays a miss). Written for issue and latency: the L4, chunk, tile and brick words of
// the current voxel (every level's index is a function of the voxel coordinates alone and
// every level is stored dense, so each fetch is in bounds) are fetched together, one memory
// round trip per iteration, and the level the walk stops at, the cell it leaves and the
// phase are picked with selects instead of a chain of divergent branches; only the reads a
// decision needs (the LOD probes' topmost material, a set voxel's material) stay behind
// branches. Keeping the words of the cells the walk has not left in registers and fetching
// only on a change, or testing level by level with branches, is 6-30 % slower on the probe
// paths and batched queries and no faster on the coherent frame. FOLD = false compiles the
// slot folding out, for worlds whose window origin folds no axis (fewer registers).
// MATSPEC = true also loads the current voxel's material with those words, so a set voxel
// costs one round trip instead of two (for latency-bound launches; the load is wasted on
// every other step).
template <bool COUNT, bool FOLD = true, bool MATSPEC = false>
__device__ __forceinline__ bool walk_iter(Walk &r, WalkPhase &q, const DevWorld &w, float tmax,
                                          uint32_t flags, bool lod, HitRec &h, uint32_t *cnt) {
    if (r.t_cur > tmax) return true;
    // out of the window, or above every occupied voxel and not moving down (a miss either way)
    const int32_t ylim = r.step[1] >= 0 ? w.ytop : w.dims[1];
    if ((uint32_t)r.l[0] >= (uint32_t)w.dims[0] || (uint32_t)r.l[1] >= (uint32_t)ylim ||
        (uint32_t)r.l[2] >= (uint32_t)w.dims[2])
        return true;
    int32_t sl[3] = {r.l[0], r.l[1], r.l[2]};
    if (FOLD && w.fold) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int32_t s = r.l[k] + w.ro_mod[k];
            sl[k] = s >= w.dims[k] ? s - w.dims[k] : s;
        }
    }
    const uint32_t l4i = l4_index(w, sl);
    const uint32_t ci = chunk_index(w, sl);
    const int cbc = cbit((sl[0] >> 6) & 3, (sl[1] >> 6) & 3, (sl[2] >> 6) & 3);
    const int tb = cbit((sl[0] >> 4) & 3, (sl[1] >> 4) & 3, (sl[2] >> 4) & 3);
    const int bb = cbit((sl[0] >> 2) & 3, (sl[1] >> 2) & 3, (sl[2] >> 2) & 3);
    const int vb = cbit(sl[0] & 3, sl[1] & 3, sl[2] & 3);
    const uint32_t tn = ci * 64u + (uint32_t)tb;
    const uint32_t bn = tn * 64u + (uint32_t)bb;
    // every level is stored dense, so each index is in bounds; one round trip for all six
    const uint64_t l4w = w.l4[l4i];
    const uint64_t cw = w.chunk[ci];
    const uint64_t tw = w.tile[tn];
    const uint32_t tu = w.tile_uni[tn];
    const uint64_t ow = w.occ[bn];
    const uint32_t bu = w.brick_uni[bn];
    const uint32_t vms = MATSPEC ? (uint32_t)w.mat[(uint64_t)bn * 64u + (uint32_t)vb] : 0u;
    asm volatile("" ::"v"(l4w), "v"(cw), "v"(tw), "v"(tu), "v"(ow), "v"(bu), "v"(vms));
    const bool p1 = q.phase1;
    // the levels the walk reaches (each test only counts if the ones above it passed)
    const bool c_on = l4w != 0 && ((l4w >> cbc) & 1);
    const bool t_on = c_on && ((cw >> tb) & 1);
    if (t_on && lod && !p1 && r.t_cur > TILE_LOD_T) {
        if (COUNT) cnt[VG_IT_TILE_LOD]++;
        if (tw) {
            const uint32_t fb = tn * 64u + (uint32_t)__builtin_ctzll(tw);
            const uint32_t lm = topmost_material(w, fb, w.occ[fb]);
            if (lm != 0u && uniform_optimisable(lm)) {
                emit(r, w, lm, h);
                return true;
            }
        }
    }
    const bool tu_on = t_on && tu != 0u;
    const bool b_on = t_on && tu == 0u && ((tw >> bb) & 1);
    const bool bu_on = b_on && bu != 0u;
    const bool v_lvl = b_on && bu == 0u;
    if (v_lvl && lod && !p1 && r.t_cur > LOD_BRICK_T) {
        if (COUNT) cnt[VG_IT_BRICK_LOD]++;
        const uint32_t m = topmost_material(w, bn, ow);
        if (!is_foliage(m)) {
            emit(r, w, m, h);
            return true;
        }
    }
    const bool v_set = v_lvl && ((ow >> vb) & 1);
    uint32_t vm = 0u;
    if (v_set) vm = MATSPEC ? vms : (uint32_t)w.mat[(uint64_t)bn * 64u + (uint32_t)vb];
    // a uniform tile, a uniform brick or a set voxel decides by its material
    const uint32_t um = tu_on ? (uint32_t)tu : (bu_on ? (uint32_t)bu : vm);
    const bool solid = (tu_on || bu_on || v_set) && solid_mat(um, flags);
    if (COUNT) {
        int cls;
        if (l4w == 0) cls = VG_IT_L4_SKIP;
        else if (!c_on) cls = !(l4w & group_mask(sl, 7)) ? VG_IT_CHUNK_GROUP_SKIP : VG_IT_CHUNK_SKIP;
        else if (!t_on) cls = !(cw & group_mask(sl, 5)) ? VG_IT_TILE_GROUP_SKIP : VG_IT_TILE_SKIP;
        else if (tu_on) cls = VG_IT_TILE_UNIFORM;
        else if (!b_on)
            cls = (!(tw & group_mask(sl, 3)) && (!lod || r.t_cur + 14.0f <= TILE_LOD_T))
                      ? VG_IT_BRICK_GROUP_SKIP
                      : VG_IT_BRICK_SKIP;
        else if (bu_on) cls = VG_IT_BRICK_UNIFORM;
        else if (v_set) cls = VG_IT_VOXEL;
        else
            cls = (!(ow & group_mask(sl, 1)) && (!lod || r.t_cur + 4.0f <= LOD_BRICK_T))
                      ? VG_IT_VOXEL_GROUP_SKIP
                      : VG_IT_VOXEL;
        cnt[cls]++;
    }
    if (solid && !p1) {
        emit(r, w, um, h);
        return true;
    }
    // the cell to leave: the empty cell (or 2x2x2 group of empty children) at the level the
    // walk stopped, the uniform cell, or the voxel (a phase-1 solid keeps the phase)
    const bool g7 = !(l4w & group_mask(sl, 7));
    const bool g5 = !(cw & group_mask(sl, 5));
    const bool g3 = !(tw & group_mask(sl, 3)) && (!lod || r.t_cur + 14.0f <= TILE_LOD_T);
    const bool g1 = !(ow & group_mask(sl, 1)) && (!lod || r.t_cur + 4.0f <= LOD_BRICK_T);
    const int32_t S = l4w == 0 ? 256
                      : !c_on    ? (g7 ? 128 : 64)
                      : !t_on    ? (g5 ? 32 : 16)
                      : tu_on    ? 16
                      : !b_on    ? (g3 ? 8 : 4)
                      : bu_on    ? 4
                      : v_set    ? 1
                                 : (g1 ? 2 : 1);
    q.phase1 = p1 && solid;
    walk_advance(r, sl, S);
    return false;
}

template <bool COUNT>
__device__ __forceinline__ void walk_run(Walk &r, const DevWorld &w, float tmax, uint32_t flags,
                                         HitRec &h, uint32_t *cnt) {
    const bool lod = (flags & VG_RAY_LOD) && !(flags & (VG_RAY_SKIP_TRANSPARENT | VG_RAY_ANY_HIT));
    WalkPhase q;
    phase_reset(q);
    // every iteration crosses at least one voxel plane inside the window
    const int max_iter = w.dims[0] + w.dims[1] + w.dims[2] + 8;
    for (int it = 0; it < max_iter; ++it)
        if (walk_iter<COUNT>(r, q, w, tmax, flags, lod, h, cnt)) return;
}

// glam normalize_or_zero: d * (1 / sqrt(dot)); false when that is not a finite positive
// reciprocal or the result is zero.
__device__ __forceinline__ bool normalize_or_zero(const float in[3], float out[3]) {
    float len = sqrtf(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
    float rcp = 1.0f / len;
    if (!(isfinite(rcp) && rcp > 0.0f)) return false;
    out[0] = in[0] * rcp;
    out[1] = in[1] * rcp;
    out[2] = in[2] * rcp;
    return !(out[0] == 0.0f && out[1] == 0.0f && out[2] == 0.0f);
}

// Pinhole direction (shaders/raymarch.wgsl:430-438, 600).
__device__ __forceinline__ void primary_dir(const CamParams &c, int32_t px, int32_t py,
                                            float out[3]) {
    float u = ((float)px + 0.5f + c.jitter[0]) / c.res[0];
    float v = ((float)py + 0.5f + c.jitter[1]) / c.res[1];
    float nx = u * 2.0f - 1.0f, ny = 1.0f - v * 2.0f;
    float aspect = c.res[0] / c.res[1];
    float a = nx * c.tan_half_fov * aspect, b = ny * c.tan_half_fov;
    float r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = c.forward[k] + c.right[k] * a + c.up[k] * b;
    float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    float rcp = 1.0f / len;
    out[0] = r[0] * rcp;
    out[1] = r[1] * rcp;
    out[2] = r[2] * rcp;
}

// p_hit + n * 0.001 (shaders/raymarch.wgsl:3955).
__device__ __forceinline__ void shadow_origin(const float o[3], const float d[3], const HitRec &h,
                                              float out[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float n = 0.0f;
        if (h.face < 6u && (int)(h.face >> 1) == k) n = (h.face & 1u) ? -1.0f : 1.0f;
        float p = o[k] + d[k] * h.t;
        out[k] = p + n * 0.001f;
    }
}

__device__ __forceinline__ void store_hit(vg_hit *dst, const HitRec &h) {
    uint32_t *p = reinterpret_cast<uint32_t *>(dst);
    p[0] = (uint32_t)h.v[0];
    p[1] = (uint32_t)h.v[1];
    p[2] = (uint32_t)h.v[2];
    p[3] = __float_as_uint(h.t);
    p[4] = (h.face & 0xFFu) | ((h.mat & 0xFFu) << 8) | ((h.hit & 0xFFu) << 16);
}

__device__ __forceinline__ void load_hit(const vg_hit *src, HitRec &h) {
    const uint32_t *p = reinterpret_cast<const uint32_t *>(src);
    h.v[0] = (int32_t)p[0];
    h.v[1] = (int32_t)p[1];
    h.v[2] = (int32_t)p[2];
    h.t = __uint_as_float(p[3]);
    h.face = p[4] & 0xFFu;
    h.mat = (p[4] >> 8) & 0xFFu;
    h.hit = (p[4] >> 16) & 0xFFu;
}

// Wave-reduce the per-lane counters and add them to the global totals.
__device__ __forceinline__ void flush_counts(const uint32_t *cnt, unsigned long long *counts) {
#pragma unroll
    for (int k = 0; k < VG_IT_COUNT; ++k) {
        unsigned long long v = cnt[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&counts[k], v);
    }
}


constexpr int kBlock = 256;

// Spread consecutive launch-order blocks round-robin over the 8 XCDs back into
// contiguous ranges per XCD, so neighbouring screen blocks share one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nb) {
    uint32_t full = nb & ~7u;
    if (bid >= full) return bid;
    return (bid & 7u) * (full >> 3) + (bid >> 3);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-local reservation of work ids for the persistent batched-query and probe path
// kernels: a refill hands its idle lanes the next ids of the reservation and takes a new
// chunk of ids (kIdChunk, or a tuned size) from the global counter (one atomic per chunk,
// not per refill) only when the reservation runs short. Single-address atomics serialise across the chip, so fewer of
// them lets waves refill more often. (The frame kernel keeps one atomic per refill: its
// refills must come from the front of the screen order to stay coherent.) All members are
// wave-uniform.
constexpr uint32_t kIdChunk = 128;
struct IdQueue {
    uint32_t next, end;
    bool drained; // the global counter has passed `total`
};

__device__ __forceinline__ void queue_reset(IdQueue &iq) {
    iq.next = iq.end = 0u;
    iq.drained = false;
}

__device__ __forceinline__ bool queue_empty(const IdQueue &iq) {
    return iq.drained && iq.next >= iq.end;
}

// Ids for `need` (<= 64) idle lanes: the idle lane of rank `rank` gets the returned id when
// `want` is set (~0u: none left). A short reservation takes max(chunk, shortfall) ids with
// one atomic (chunk 0: exactly what the refill lacks, no reservation left over). Every lane
// of the wave must call it together.
__device__ __forceinline__ uint32_t queue_take(IdQueue &iq, uint32_t *work, uint32_t total,
                                               uint32_t need, uint32_t rank, bool want,
                                               uint32_t lane, uint32_t chunk = kIdChunk) {
    const uint32_t avail = iq.end - iq.next;
    uint32_t b = ~0u, g = 0u;
    if (avail < need && !iq.drained) {
        g = need - avail > chunk ? need - avail : chunk;
        if (lane == 0) b = atomicAdd(work, g);
        b = __shfl(b, 0, 64);
        if (b >= total) {
            iq.drained = true;
            b = ~0u;
        }
    }
    uint32_t id = ~0u;
    if (want) {
        if (rank < avail) id = iq.next + rank;
        else if (b != ~0u) id = b + (rank - avail);
        if (id >= total) id = ~0u;
    }
    if (b != ~0u) {
        const uint32_t end = total - b < g ? total : b + g;
        const uint32_t nx = b + (need - avail);
        iq.next = nx < end ? nx : end;
        iq.end = end;
    } else {
        iq.next += avail < need ? avail : need;
    }
    return id;
}

} // namespace vg
