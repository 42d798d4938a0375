// Device engine for gfx950 (MI355X, CDNA4, wave64): world storage in HBM, the
// relayout kernels of the upload path, and the exact hierarchical voxel walk used
// by every trace entry point.
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
#include <mutex>
#include <vector>

#include "vg_world.h"

namespace vg {

// Fused beam prologue (the role of cs_beam, shaders/beam.wgsl:31-113, made conservative).
// A wave is one 8x8 pixel block whose 64 rays share the camera origin O, so on [t0, t1]
// they all lie in O + [min d, max d] * [t0, t1] (per axis, min / max over the wave). The
// wave marches that box in 16-voxel steps, grown by one voxel, each lane testing one world
// tile the box touches inside the window (a 4x4x4 cube at most) against the chunk masks; the block
// starts at the first step whose box touches an occupied tile. Every voxel a block ray
// visits before that t is empty, so walk_advance_to may jump there exactly. Every lane of
// the wave must call this (it reduces and votes across the wave).
constexpr float kBeamStep = 16.0f;
constexpr size_t kBeamDefaultBlocks = 512 * 512; // 8x8 blocks of a 4096x4096 image
constexpr int kBeamMaxSteps = 64;

template <bool COUNT>
__device__ __forceinline__ float beam_start(const DevWorld &w, const CamParams &c,
                                            const float d[3], float tmax, uint32_t *cnt) {
    bool inside = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float rel = c.origin[k] - (float)w.origin[k];
        inside = inside && rel >= 0.0f && rel < (float)w.dims[k];
    }
    if (!inside) return 0.0f; // wave-uniform: the camera is shared
    float dmin[3], dmax[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = d[k], hi = d[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo = fsel_min(lo, __shfl_xor(lo, off, 64));
            hi = fsel_max(hi, __shfl_xor(hi, off, 64));
        }
        dmin[k] = lo;
        dmax[k] = hi;
    }
    const int lane = threadIdx.x & 63;
    // every block ray moves up (or level) and the step's box lies at or above the world's
    // occupied y bound: nothing is left to hit, as when the box leaves the window
    const bool rising = dmin[1] >= 0.0f;
    const float ybound = (float)(w.origin[1] + w.ytop);
    // Packed rounds: 8 consecutive steps at once, 8 lanes each, while every box fits 8 tiles
    // (the common case: a block's box is 1-2 tiles wide per axis); the first step that stops
    // the march decides, exactly as the one-step loop below would. A step whose box needs
    // more than 8 tiles hands over to that loop.
    int st = 0;
    {
        const int g = lane >> 3, j = lane & 7;
        for (;;) {
            const int s = st + g;
            const float ts0 = (float)s * kBeamStep, ts1 = (float)(s + 1) * kBeamStep;
            // 0 test, 1 stop before (loop bound), 2 exit (box left the window), 3 too wide
            // (stop), 4 more than 8 tiles (hand over)
            int kind = 0;
            int32_t lo[3], n[3];
            int32_t total = 1;
            if (s >= kBeamMaxSteps || !(ts0 < tmax)) {
                kind = 1;
            } else {
                float mny = 0.0f;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    float a = c.origin[k] + dmin[k] * ts0, b = c.origin[k] + dmin[k] * ts1;
                    float cc = c.origin[k] + dmax[k] * ts0, e = c.origin[k] + dmax[k] * ts1;
                    float mn = fsel_min(fsel_min(a, b), fsel_min(cc, e)) - 1.0f;
                    float mx = fsel_max(fsel_max(a, b), fsel_max(cc, e)) + 1.0f;
                    if (k == 1) mny = mn;
                    int32_t tl = (int32_t)floorf(mn * 0.0625f), th = (int32_t)floorf(mx * 0.0625f);
                    const int32_t wl = w.origin[k] >> 4, wh = (w.origin[k] + w.dims[k] - 1) >> 4;
                    tl = tl < wl ? wl : tl;
                    th = th > wh ? wh : th;
                    lo[k] = tl;
                    n[k] = th >= tl ? th - tl + 1 : 0;
                    total *= n[k];
                }
                if (total == 0 || (rising && mny >= ybound)) kind = 2;
                else if (n[0] > 4 || n[1] > 4 || n[2] > 4) kind = 3;
                else if (total > 8) kind = 4;
            }
            bool occ = false;
            const bool tests = kind == 0 && j < total;
            if (tests) {
                // j < 8 over a box of n[0] x n[1] x n[2] <= 8 tiles, each n <= 4: the
                // quotients come from a float reciprocal ((j + 0.5) / m lies at least 1/16
                // from any integer for m <= 8, far beyond v_rcp_f32's error), not from
                // integer division
                const int32_t n01 = n[0] * n[1];
                const int32_t q0 = (int32_t)(((float)j + 0.5f) * __builtin_amdgcn_rcpf((float)n[0]));
                const int32_t q1 = (int32_t)(((float)q0 + 0.5f) * __builtin_amdgcn_rcpf((float)n[1]));
                const int32_t q2 = (int32_t)(((float)j + 0.5f) * __builtin_amdgcn_rcpf((float)n01));
                const int32_t tw[3] = {lo[0] + (j - q0 * n[0]), lo[1] + (q0 - q1 * n[1]),
                                       lo[2] + q2};
                int32_t sl[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    int32_t v = tw[k] * 16 - w.origin[k] + w.ro_mod[k];
                    v = v < 0 ? v + w.dims[k] : v;
                    sl[k] = v >= w.dims[k] ? v - w.dims[k] : v;
                }
                const uint32_t ci = chunk_index(w, sl);
                occ = (w.chunk[ci] >> cbit((sl[0] >> 4) & 3, (sl[1] >> 4) & 3, (sl[2] >> 4) & 3)) & 1;
            }
            // the first step (group) that stops the march
            const uint64_t stop = __ballot(kind != 0 || occ);
            if (!stop) {
                if (COUNT && tests) cnt[VG_IT_BEAM_TEST]++;
                st += 8;
                continue;
            }
            const int gs = (__ffsll((long long)stop) - 1) >> 3;
            const int kind_s = __shfl(kind, gs * 8, 64);
            if (COUNT && tests && (g < gs || (g == gs && kind_s == 0))) cnt[VG_IT_BEAM_TEST]++;
            const float t_s = (float)(st + gs) * kBeamStep;
            if (kind_s == 4) {
                st += gs; // this step needs the wide loop
                break;
            }
            return kind_s == 2 ? t_s + kBeamStep : t_s;
        }
    }
    float t0 = (float)st * kBeamStep;
    for (; st < kBeamMaxSteps && t0 < tmax; ++st) {
        const float t1 = (float)(st + 1) * kBeamStep;
        int32_t lo[3], n[3];
        int32_t total = 1;
        float mny = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float a = c.origin[k] + dmin[k] * t0, b = c.origin[k] + dmin[k] * t1;
            float cc = c.origin[k] + dmax[k] * t0, e = c.origin[k] + dmax[k] * t1;
            float mn = fsel_min(fsel_min(a, b), fsel_min(cc, e)) - 1.0f;
            float mx = fsel_max(fsel_max(a, b), fsel_max(cc, e)) + 1.0f;
            if (k == 1) mny = mn;
            int32_t tl = (int32_t)floorf(mn * 0.0625f), th = (int32_t)floorf(mx * 0.0625f);
            const int32_t wl = w.origin[k] >> 4, wh = (w.origin[k] + w.dims[k] - 1) >> 4;
            tl = tl < wl ? wl : tl;
            th = th > wh ? wh : th;
            lo[k] = tl;
            n[k] = th >= tl ? th - tl + 1 : 0;
            total *= n[k];
        }
        if (total == 0 || (rising && mny >= ybound)) {
            // the box left the window: every block ray has exited it for good (the window is
            // convex), so nothing of the window remains to visit; or every block ray rises
            // above every occupied voxel from here on
            t0 = t1;
            break;
        }
        if (n[0] > 4 || n[1] > 4 || n[2] > 4) break; // wider than the wave's 4x4x4 tile cube
        // lane -> one tile of a 4x4x4 cube at lo
        const int32_t ix = lane & 3, iz = (lane >> 2) & 3, iy = lane >> 4;
        bool occ = false;
        if (ix < n[0] && iy < n[1] && iz < n[2]) {
            if (COUNT) cnt[VG_IT_BEAM_TEST]++;
            const int32_t tw[3] = {lo[0] + ix, lo[1] + iy, lo[2] + iz};
            int32_t sl[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                // tile start, window-local (>= -15), folded into its slot
                int32_t v = tw[k] * 16 - w.origin[k] + w.ro_mod[k];
                v = v < 0 ? v + w.dims[k] : v;
                sl[k] = v >= w.dims[k] ? v - w.dims[k] : v;
            }
            const uint32_t ci = chunk_index(w, sl);
            occ = (w.chunk[ci] >> cbit((sl[0] >> 4) & 3, (sl[1] >> 4) & 3, (sl[2] >> 4) & 3)) & 1;
        }
        if (__ballot(occ) != 0ull) break;
        t0 = t1;
    }
    return t0;
}

// ------------------------------------------------------------------ kernels --

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_trace_rays(DevWorld w, const vg_ray *__restrict__ rays,
                                                       int64_t n, uint32_t flags,
                                                       vg_hit *__restrict__ hits,
                                                       unsigned long long *counts) {
    world_ready(w);
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t cnt[VG_IT_COUNT];
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    if (i < n) {
        const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
        float4 a = rp[0], b = rp[1];
        float o[3] = {a.x, a.y, a.z}, din[3] = {b.x, b.y, b.z}, d[3];
        uint32_t f = flags | __float_as_uint(b.w);
        HitRec h;
        hit_miss(h);
        if (normalize_or_zero(din, d)) {
            Walk r;
            if (walk_init(r, w, o, d)) walk_run<COUNT>(r, w, a.w, f, h, cnt);
        }
        store_hit(hits + i, h);
    }
    if (COUNT) flush_counts(cnt, counts);
}

// One thread per pixel, a wave per 8x8 pixel block, 16x16 pixels per workgroup.
template <bool COUNT, int MODE, bool BEAM> // MODE 0 primary, 1 primary+shadow, 2 shadow from hits
__global__ __launch_bounds__(kBlock) void k_trace_frame(DevWorld w, CamParams cam, int32_t width,
                                                        int32_t height, uint32_t nbx,
                                                        float ptmax, uint32_t pflags, float sx,
                                                        float sy, float sz, float stmax,
                                                        vg_hit *__restrict__ hits,
                                                        uint8_t *__restrict__ occluded,
                                                        unsigned long long *counts) {
    world_ready(w);
    const uint32_t nb = gridDim.x;
    const uint32_t b = xcd_block(blockIdx.x, nb);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t px = (int32_t)((b % nbx) * 16 + (wave & 1) * 8 + (lane & 7));
    const int32_t py = (int32_t)((b / nbx) * 16 + (wave >> 1) * 8 + (lane >> 3));
    uint32_t cnt[VG_IT_COUNT];
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    float d[3];
    primary_dir(cam, px, py, d); // every lane, also past the image edge (the beam reads all)
    float tb = 0.0f;
    if (BEAM && MODE != 2) tb = beam_start<COUNT>(w, cam, d, ptmax, cnt);
    if (px < width && py < height) {
        const int64_t i = (int64_t)py * width + px;
        HitRec h;
        if (MODE == 2) {
            load_hit(hits + i, h);
        } else {
            hit_miss(h);
            Walk r;
            if (walk_init(r, w, cam.origin, d)) {
                if (BEAM && tb > 0.0f && r.t_cur <= ptmax) walk_advance_to(r, tb);
                walk_run<COUNT>(r, w, ptmax, pflags, h, cnt);
            }
            store_hit(hits + i, h);
        }
        if (MODE != 0) {
            uint32_t occ = 0;
            if (h.hit) {
                float so[3], sd[3] = {sx, sy, sz};
                shadow_origin(cam.origin, d, h, so);
                HitRec hs;
                hit_miss(hs);
                Walk r2;
                if (walk_init(r2, w, so, sd)) walk_run<COUNT>(r2, w, stmax, VG_RAY_ANY_HIT, hs, cnt);
                occ = hs.hit;
            }
            occluded[i] = (uint8_t)occ;
        }
    }
    if (COUNT) flush_counts(cnt, counts);
}


// Striped work counters of the persistent frame kernel: stripe c's counter is work[c *
// kStripeWords], one 128-byte line each.
constexpr uint32_t kStripeWords = 32;
constexpr uint32_t kMaxStripes = 64;
constexpr size_t kStripeBytes = kMaxStripes * kStripeWords * sizeof(uint32_t);

// Parameters of a frame launch that only the persistent kernel's refills and shadow-ray
// set-ups read. The launch's first kernel (k_beam or k_frame_setup) writes them from its own
// arguments to device memory after the counters, and the persistent kernel loads them where
// it needs them, so they do not hold scalar registers through the walk loop (held there,
// they spilled the world's pointers into vector lanes, reloaded on every pyramid fetch).
struct FrameParams {
    CamParams cam;
    float sun[3];
    float sun_inv[3]; // safe_inv of each sun component, the shadow rays' shared inverse
    int32_t width, height;
    uint32_t nbx;   // pixel blocks per image row
    uint32_t bws;   // log2 of the block width: blocks of 2^bws x 64 / 2^bws pixels (a wave)
    // the persistent launch's static first reservations (frame_static): its workgroups, its
    // stripes, the frame's pixel blocks (0 for a block list: no static reservations)
    uint32_t grid_wg, stripes, nblk;
};
constexpr size_t kWorkBytes = kStripeBytes + 256; // counters, then the FrameParams
static_assert(sizeof(FrameParams) <= 256, "FrameParams must fit its slot");

__device__ __forceinline__ FrameParams *frame_params(uint32_t *work) {
    return reinterpret_cast<FrameParams *>(reinterpret_cast<uint8_t *>(work) + kStripeBytes);
}

// Ids [base, lim) of stripe sc for `need` idle lanes (one atomic on the stripe's counter); a
// wave whose stripe is used up moves on to the next stripe, and base = ~0u once it has found
// every stripe used up. Wave-uniform: every lane calls it together. (The counters are the
// only shared state: flags read with agent-scope atomic loads beside them cost the kernel
// ~30 % on this part, measured.)
__device__ __forceinline__ void stripe_take(uint32_t *work, uint32_t stripes, uint32_t nblk,
                                            uint32_t sblk, uint32_t need, uint32_t lane,
                                            uint32_t &sc, uint32_t &base, uint32_t &lim) {
    // stripe c: blocks [c sblk, min((c + 1) sblk, nblk)), sblk = ceil(nblk / stripes)
    base = ~0u;
    for (uint32_t tries = 0; tries < stripes; ++tries) {
        const uint32_t b0 = sc * sblk < nblk ? sc * sblk : nblk;
        const uint32_t b1 = b0 + sblk < nblk ? b0 + sblk : nblk;
        const uint32_t s0 = b0 * 64u, s1 = b1 * 64u;
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(work + sc * kStripeWords, need);
        // wave-uniform from here on: scalar registers and branches
        b = __builtin_amdgcn_readfirstlane(__shfl(b, 0, 64));
        if (b < s1 - s0) {
            base = s0 + b;
            lim = s1 - base < need ? s1 : base + need;
            return;
        }
        sc = sc + 1u < stripes ? sc + 1u : 0u;
    }
}

// Zero the striped counters (one thread per stripe).
// Static first reservations of the persistent frame kernel. Workgroup b starts on stripe
// frame_stripe(b) (round-robin over the XCDs, as the kernel picks it); the waves of the
// workgroups of one stripe, in the order of b, take its first pixel blocks one each, so no
// wave waits on the stripe's atomic for its first 64 rays (thousands of waves on a few
// counters at the launch's start take ~20 us); the counters start past those blocks.
__device__ __forceinline__ uint32_t frame_stripe(uint32_t b, uint32_t stripes) {
    const uint32_t per_xcd = stripes >= 8u ? stripes / 8u : 1u;
    return stripes >= 8u ? (b % 8u) * per_xcd + (b / 8u) % per_xcd : b % stripes;
}
// the workgroups before b on b's stripe
__device__ __forceinline__ uint32_t frame_stripe_rank(uint32_t b, uint32_t stripes) {
    return stripes >= 8u ? (b / 8u) / (stripes / 8u) : b / stripes;
}
// the workgroups of a grid_wg-workgroup launch on stripe sc
__device__ __forceinline__ uint32_t frame_stripe_wgs(uint32_t sc, uint32_t stripes,
                                                     uint32_t grid_wg) {
    if (stripes >= 8u) {
        const uint32_t per_xcd = stripes / 8u;
        if (sc >= 8u * per_xcd) return 0u;
        const uint32_t first = sc / per_xcd + 8u * (sc % per_xcd); // its lowest workgroup
        return first < grid_wg ? (grid_wg - 1u - first) / (8u * per_xcd) + 1u : 0u;
    }
    return sc < grid_wg ? (grid_wg - 1u - sc) / stripes + 1u : 0u;
}
// pixel blocks [b0, b1) of stripe sc: sblk = ceil(nblk / stripes) each
__device__ __forceinline__ void frame_stripe_blocks(uint32_t sc, uint32_t sblk, uint32_t nblk,
                                                    uint32_t &b0, uint32_t &b1) {
    b0 = sc * sblk < nblk ? sc * sblk : nblk;
    b1 = b0 + sblk < nblk ? b0 + sblk : nblk;
}

// Each stripe counter (one thread per stripe) starts past the blocks its waves take statically.
__device__ __forceinline__ void stripe_reset(uint32_t *work, uint32_t t, const FrameParams &fp) {
    if (t >= kMaxStripes) return;
    uint32_t v = 0u;
    if (fp.nblk && t < fp.stripes) {
        const uint32_t sblk = (fp.nblk + fp.stripes - 1u) / fp.stripes;
        uint32_t b0, b1;
        frame_stripe_blocks(t, sblk, fp.nblk, b0, b1);
        const uint32_t waves = frame_stripe_wgs(t, fp.stripes, fp.grid_wg) * (kBlock / 64);
        v = (waves < b1 - b0 ? waves : b1 - b0) * 64u;
    }
    work[t * kStripeWords] = v;
}

// Beam pre-pass of the persistent frame: one wave per 8x8 pixel block writes the block's
// exact start t (beam_start). Its first workgroup also zeroes the frame's striped work
// counters, which the persistent kernel enqueued after it on the same stream consumes.
// With a block list (list / list_n on the device), the waves take the listed blocks only.
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_beam(DevWorld w, FrameParams fp, uint32_t nblocks,
                                                 float ptmax, float *__restrict__ beam_t,
                                                 uint32_t *work,
                                                 const uint32_t *__restrict__ list,
                                                 const uint32_t *list_n,
                                                 unsigned long long *counts) {
    world_ready(w);
    if (blockIdx.x == 0) {
        stripe_reset(work, threadIdx.x, fp);
        if (threadIdx.x == 0) *frame_params(work) = fp;
    }
    const CamParams &cam = fp.cam;
    const uint32_t nbx = fp.nbx, bws = fp.bws;
    const uint32_t wi = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (wi >= (list ? *list_n : nblocks)) return; // whole waves leave together
    const uint32_t blk = list ? list[wi] : wi;
    const uint32_t lane = threadIdx.x & 63;
    const int32_t px = (int32_t)(((blk % nbx) << bws) + (lane & ((1u << bws) - 1u)));
    const int32_t py = (int32_t)((blk / nbx) * (64u >> bws) + (lane >> bws));
    uint32_t cnt[VG_IT_COUNT];
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    float d[3];
    primary_dir(cam, px, py, d);
    const float tb = beam_start<COUNT>(w, cam, d, ptmax, cnt);
    if (lane == 0) beam_t[blk] = tb;
    if (COUNT) flush_counts(cnt, counts);
}

__global__ void k_reset_work(uint32_t *work) { *work = 0u; }

// The frame launch's first kernel without the beam: counters and FrameParams.
__global__ void k_frame_setup(uint32_t *work, FrameParams fp) {
    stripe_reset(work, threadIdx.x, fp);
    if (threadIdx.x == 0) *frame_params(work) = fp;
}

// Persistent, ray-compacting frame kernel. Every wave keeps its 64 lanes busy: whenever
// fewer than `threshold` lanes hold a live ray, the wave takes as many new primary rays as
// it has idle lanes from one global counter (one atomic per wave) and hands them out by
// lane rank. Ray ids run over 8x8 pixel blocks, so a refill stays spatially coherent. A
// primary hit becomes that lane's shadow ray at once. The grid is sized to the machine
// (waves per CU chosen by the autotuner), not to the image. FOLD = false: the world's window
// folds no axis (the host picks the instantiation).
template <bool COUNT, bool SHADOW, bool BEAM, bool FOLD>
__global__ __launch_bounds__(kBlock) void k_frame_persistent(
    DevWorld w, uint32_t total_ids, float ptmax, uint32_t pflags, float stmax,
    const float *__restrict__ beam_t, vg_hit *__restrict__ hits, uint8_t *__restrict__ occluded,
    uint32_t *work, int threshold, uint32_t stripes, const uint32_t *__restrict__ list,
    const uint32_t *list_n, unsigned long long *counts) {
    world_ready(w);
    const FrameParams *__restrict__ fpp = frame_params(work);
    // with a block list, ray ids run over the listed 8x8 blocks only
    const uint32_t lane = threadIdx.x & 63;
    if (list) total_ids = *list_n * 64u;
    // striped work queue: the 8x8 blocks are cut into `stripes` contiguous stripes, each with
    // its own counter; a wave starts on a stripe of its XCD (workgroups go round-robin over
    // the 8 XCDs) and moves on to the next open stripe when its own is used up
    const uint32_t nblk = total_ids / 64u;
    const uint32_t sblk = (nblk + stripes - 1u) / stripes;
    const uint32_t per_xcd = stripes >= 8u ? stripes / 8u : 1u;
    uint32_t sc = stripes >= 8u ? (blockIdx.x % 8u) * per_xcd + (blockIdx.x / 8u) % per_xcd
                                : blockIdx.x % stripes;
    // the wave's static first reservation: one pixel block of its stripe (frame_stripe)
    uint32_t st_base = 0u, st_lim = 0u;
    if (!list) {
        uint32_t b0, b1;
        frame_stripe_blocks(sc, sblk, nblk, b0, b1);
        const uint32_t blk = b0 + frame_stripe_rank(blockIdx.x, stripes) * (kBlock / 64) +
                             (threadIdx.x >> 6);
        if (blk < b1) {
            st_base = blk * 64u;
            st_lim = st_base + 64u;
        }
    }
    uint32_t cnt[VG_IT_COUNT];
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    const bool plod =
        (pflags & VG_RAY_LOD) && !(pflags & (VG_RAY_SKIP_TRANSPARENT | VG_RAY_ANY_HIT));
    const int max_iter = w.dims[0] + w.dims[1] + w.dims[2] + 8;
    int kind = 0; // 0 idle, 1 primary ray, 2 shadow ray
    int64_t pix = 0;
    int it = 0;
    Walk r;
    WalkPhase q;
    HitRec h;
    hit_miss(h);
    phase_reset(q);
    bool exhausted = false;
    for (;;) {
        uint64_t act = __ballot(kind != 0);
        int nact = __popcll(act);
        if (!exhausted && nact < threshold) {
            uint32_t base = ~0u, lim = 0u; // ids [base, lim) for the idle lanes
            if (st_lim) { // the first refill (every lane idle): the static reservation
                base = st_base;
                lim = st_lim;
                st_lim = 0u;
            } else {
                stripe_take(work, stripes, nblk, sblk, 64u - (uint32_t)nact, lane, sc, base,
                            lim);
            }
            if (base == ~0u) {
                exhausted = true;
            } else if (kind == 0) {
                const uint32_t id = base + lanes_below(~act);
                if (id < lim) {
                    const uint32_t nbx = fpp->nbx, bws = fpp->bws;
                    const int32_t width = fpp->width, height = fpp->height;
                    const uint32_t blk = list ? list[id >> 6] : id >> 6, qd = id & 63u;
                    const int32_t px = (int32_t)(((blk % nbx) << bws) + (qd & ((1u << bws) - 1u)));
                    const int32_t py = (int32_t)((blk / nbx) * (64u >> bws) + (qd >> bws));
                    if (px < width && py < height) {
                        pix = (int64_t)py * width + px;
                        const CamParams cam = fpp->cam;
                        float d[3];
                        primary_dir(cam, px, py, d);
                        hit_miss(h);
                        if (walk_init(r, w, cam.origin, d)) {
                            if (BEAM) {
                                const float tb = beam_t[blk];
                                if (tb > 0.0f && r.t_cur <= ptmax) walk_advance_to(r, tb);
                            }
                            phase_reset(q);
                            it = 0;
                            kind = 1;
                        } else {
                            store_hit(hits + pix, h);
                            if (SHADOW) occluded[pix] = 0;
                        }
                    }
                }
            }
            nact = __popcll(__ballot(kind != 0));
        }
        if (nact == 0) {
            if (exhausted) break;
            continue;
        }
        do {
            if (kind != 0) {
                // one inlined walk for both kinds (per-lane parameters), so a wave holding
                // primary and shadow rays at once does not execute two walk bodies
                const bool prim = kind == 1;
                bool done = walk_iter<COUNT, FOLD>(r, q, w, prim ? ptmax : stmax,
                                                  prim ? pflags : (uint32_t)VG_RAY_ANY_HIT,
                                                  prim && plod, h, cnt);
                done = done || ++it >= max_iter;
                if (done) {
                    if (kind == 1) {
                        store_hit(hits + pix, h);
                        kind = 0;
                        if (SHADOW) {
                            bool started = false;
                            if (h.hit) {
                                // the primary walk still holds its direction, bit for bit
                                // the pinhole vector primary_dir gave it
                                const float co[3] = {fpp->cam.origin[0], fpp->cam.origin[1],
                                                     fpp->cam.origin[2]};
                                const float sd[3] = {fpp->sun[0], fpp->sun[1], fpp->sun[2]};
                                const float si[3] = {fpp->sun_inv[0], fpp->sun_inv[1],
                                                     fpp->sun_inv[2]};
                                float so[3];
                                shadow_origin(co, r.dir, h, so);
                                hit_miss(h);
                                if (walk_init(r, w, so, sd, si)) {
                                    phase_reset(q);
                                    it = 0;
                                    kind = 2;
                                    started = true;
                                }
                            }
                            if (!started) occluded[pix] = 0;
                        }
                    } else {
                        occluded[pix] = (uint8_t)h.hit;
                        kind = 0;
                    }
                }
            }
        } while (__popcll(__ballot(kind != 0)) >= threshold);
    }
    if (COUNT) flush_counts(cnt, counts);
}

// ------------------------------------------------------------ water rays --
// The secondary rays shade_water_top casts at a water surface (shaders/raymarch.wgsl:
// 3534-3667): at the primary hit p = origin + d t with face normal n, the reflection
// reflect(d, n) from p + n (0.01 + 0.34 graze^2), graze = 1 - clamp(-d . n, 0, 1), and the
// refraction refract(d, n, 1/1.33) (d itself on total internal reflection, length < 0.01)
// from p + d 0.001. Both are traced as batched queries: direction normalize_or_zero'd,
// VG_RAY_SKIP_TRANSPARENT (trace_no_water), tmax SECONDARY_MAX_T = 200.
constexpr float kSecondaryMaxT = 200.0f;

__device__ __forceinline__ bool is_water(uint32_t m) { return m >= 5u && m <= 12u; }

__device__ __forceinline__ void water_rays(const float o[3], const float d[3], float t, uint32_t face,
                                           float ro[3], float rd[3], float fo[3], float fd[3]) {
    float n[3], pw[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        n[k] = (face < 6u && (int)(face >> 1) == k) ? ((face & 1u) ? -1.0f : 1.0f) : 0.0f;
        pw[k] = o[k] + d[k] * t;
    }
    const float dn = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
    float c = -d[0] * n[0] + -d[1] * n[1] + -d[2] * n[2];
    c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    const float graze = 1.0f - c;
    const float lift = 0.01f + 0.34f * graze * graze;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rd[k] = d[k] - 2.0f * dn * n[k];
        ro[k] = pw[k] + n[k] * lift;
    }
    const float eta = (float)(1.0 / 1.33);
    const float kk = 1.0f - eta * eta * (1.0f - dn * dn);
    float f[3] = {0.0f, 0.0f, 0.0f};
    if (!(kk < 0.0f)) {
        const float sk = sqrtf(kk);
#pragma unroll
        for (int k = 0; k < 3; ++k) f[k] = eta * d[k] - (eta * dn + sk) * n[k];
    }
    const bool tir = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) < 0.01f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        fd[k] = tir ? d[k] : f[k];
        fo[k] = pw[k] + d[k] * 0.001f;
    }
}

// One thread per pixel in 8x8 block order: pixels whose primary hit is water entered through
// a face are appended to the list (one atomic per wave, so the list keeps runs of
// neighbouring pixels); every other pixel gets its two miss records here.
__global__ __launch_bounds__(kBlock) void k_water_compact(int32_t width, int32_t height, uint32_t nbx,
                                                          uint32_t total_ids,
                                                          const vg_hit *__restrict__ prim,
                                                          vg_hit *__restrict__ refl,
                                                          vg_hit *__restrict__ refr,
                                                          uint32_t *__restrict__ list,
                                                          uint32_t *list_n) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    bool water = false;
    int64_t pix = 0;
    if (id < total_ids) {
        const uint32_t blk = id >> 6, qd = id & 63u;
        const int32_t px = (int32_t)((blk % nbx) * 8 + (qd & 7));
        const int32_t py = (int32_t)((blk / nbx) * 8 + (qd >> 3));
        if (px < width && py < height) {
            pix = (int64_t)py * width + px;
            HitRec p0;
            load_hit(prim + pix, p0);
            water = p0.hit && is_water(p0.mat) && p0.face < 6u;
            if (!water) {
                HitRec m;
                hit_miss(m);
                store_hit(refl + pix, m);
                store_hit(refr + pix, m);
            }
        }
    }
    const uint64_t b = __ballot(water);
    if (!b) return; // wave-uniform
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(list_n, (uint32_t)__popcll(b));
    base = __shfl(base, 0, 64);
    if (water) list[base + lanes_below(b)] = (uint32_t)pix;
}

// Persistent, ray-compacting over the listed water pixels: a lane traces the pixel's
// reflection ray, then its refraction ray.
__global__ __launch_bounds__(kBlock) void k_water_persistent(
    DevWorld w, CamParams cam, int32_t width, const uint32_t *__restrict__ list,
    const uint32_t *list_n, const vg_hit *__restrict__ prim, vg_hit *__restrict__ refl,
    vg_hit *__restrict__ refr, uint32_t *work, int threshold) {
    world_ready(w);
    const uint32_t total_ids = *list_n;
    const uint32_t lane = threadIdx.x & 63;
    const int max_iter = w.dims[0] + w.dims[1] + w.dims[2] + 8;
    int kind = 0; // 0 idle, 1 reflection ray, 2 refraction ray
    int64_t pix = 0;
    int it = 0;
    float fo[3] = {0.0f, 0.0f, 0.0f}, fd[3] = {0.0f, 0.0f, 0.0f}; // the pending refraction ray
    Walk r;
    WalkPhase q;
    HitRec h;
    hit_miss(h);
    phase_reset(q);
    auto start = [&](const float o[3], const float draw[3], int k) -> bool {
        float d[3];
        hit_miss(h);
        if (!normalize_or_zero(draw, d) || !walk_init(r, w, o, d)) return false;
        phase_reset(q);
        it = 0;
        kind = k;
        return true;
    };
    auto refraction = [&]() {
        if (!start(fo, fd, 2)) {
            store_hit(refr + pix, h);
            kind = 0;
        }
    };
    bool exhausted = false;
    for (;;) {
        uint64_t act = __ballot(kind != 0);
        int nact = __popcll(act);
        if (!exhausted && nact < threshold) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(work, (uint32_t)(64 - nact));
            base = __shfl(base, 0, 64);
            if (base >= total_ids) {
                exhausted = true;
            } else if (kind == 0) {
                const uint32_t id = base + lanes_below(~act);
                if (id < total_ids) {
                    pix = list[id];
                    HitRec p0;
                    load_hit(prim + pix, p0);
                    float d[3], ro[3], rd[3];
                    primary_dir(cam, (int32_t)(pix % width), (int32_t)(pix / width), d);
                    water_rays(cam.origin, d, p0.t, p0.face, ro, rd, fo, fd);
                    if (!start(ro, rd, 1)) {
                        store_hit(refl + pix, h);
                        refraction();
                    }
                }
            }
            nact = __popcll(__ballot(kind != 0));
        }
        if (nact == 0) {
            if (exhausted) break;
            continue;
        }
        do {
            if (kind != 0) {
                bool done = walk_iter<false>(r, q, w, kSecondaryMaxT,
                                                   VG_RAY_SKIP_TRANSPARENT, false, h, nullptr);
                done = done || ++it >= max_iter;
                if (done) {
                    if (kind == 1) {
                        store_hit(refl + pix, h);
                        refraction();
                    } else {
                        store_hit(refr + pix, h);
                        kind = 0;
                    }
                }
            }
        } while (__popcll(__ballot(kind != 0)) >= threshold);
    }
}

// Persistent, ray-compacting batched query: the same refill discipline over a vg_ray array.
// FOLD as for k_frame_persistent.
template <bool COUNT, bool FOLD>
__global__ __launch_bounds__(kBlock) void k_rays_persistent(DevWorld w,
                                                            const vg_ray *__restrict__ rays,
                                                            uint32_t n, uint32_t flags,
                                                            vg_hit *__restrict__ hits,
                                                            uint32_t *work, int threshold,
                                                            unsigned long long *counts) {
    world_ready(w);
    const uint32_t lane = threadIdx.x & 63;
    uint32_t cnt[VG_IT_COUNT];
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    const int max_iter = w.dims[0] + w.dims[1] + w.dims[2] + 8;
    bool busy = false;
    uint32_t id = 0, f = 0;
    float tmax = 0.0f;
    bool lod = false;
    int it = 0;
    Walk r;
    WalkPhase q;
    HitRec h;
    hit_miss(h);
    phase_reset(q);
    IdQueue iq;
    queue_reset(iq);
    for (;;) {
        uint64_t act = __ballot(busy);
        int nact = __popcll(act);
        if (!queue_empty(iq) && nact < threshold) {
            const uint32_t nid = queue_take(iq, work, n, (uint32_t)(64 - nact), lanes_below(~act),
                                            !busy, lane);
            if (!busy) {
                if (nid != ~0u) {
                    id = nid;
                    const float4 *rp = reinterpret_cast<const float4 *>(rays + id);
                    const float4 a = rp[0], b = rp[1];
                    const float o[3] = {a.x, a.y, a.z}, din[3] = {b.x, b.y, b.z};
                    float d[3];
                    f = flags | __float_as_uint(b.w);
                    tmax = a.w;
                    lod = (f & VG_RAY_LOD) && !(f & (VG_RAY_SKIP_TRANSPARENT | VG_RAY_ANY_HIT));
                    hit_miss(h);
                    if (normalize_or_zero(din, d) && walk_init(r, w, o, d)) {
                        phase_reset(q);
                        it = 0;
                        busy = true;
                    } else {
                        store_hit(hits + id, h);
                    }
                }
            }
            nact = __popcll(__ballot(busy));
        }
        if (nact == 0) {
            if (queue_empty(iq)) break;
            continue;
        }
        do {
            if (busy) {
                bool done = walk_iter<COUNT, FOLD>(r, q, w, tmax, f, lod, h, cnt);
                done = done || ++it >= max_iter;
                if (done) {
                    store_hit(hits + id, h);
                    busy = false;
                }
            }
        } while (__popcll(__ballot(busy)) >= threshold);
    }
    if (COUNT) flush_counts(cnt, counts);
}

// Persistent, ray-compacting query over many uneven batches at once: ray ids run over the
// concatenation of all batches (prefix[b] = first id of batch b, prefix[nb] = total), each
// lane finds its batch by binary search, and the refill discipline is that of
// k_rays_persistent, so short and long batches share the machine without tail waste.
template <bool COUNT, bool FOLD>
__global__ __launch_bounds__(kBlock) void k_batches_persistent(DevWorld w,
                                                               const vg_batch *__restrict__ bat,
                                                               const uint32_t *__restrict__ prefix,
                                                               uint32_t nbat, uint32_t total,
                                                               uint32_t *work, int threshold,
                                                               unsigned long long *counts) {
    world_ready(w);
    const uint32_t lane = threadIdx.x & 63;
    uint32_t cnt[VG_IT_COUNT];
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    const int max_iter = w.dims[0] + w.dims[1] + w.dims[2] + 8;
    bool busy = false;
    vg_hit *out = nullptr;
    uint32_t f = 0;
    float tmax = 0.0f;
    bool lod = false;
    int it = 0;
    Walk r;
    WalkPhase q;
    HitRec h;
    hit_miss(h);
    phase_reset(q);
    IdQueue iq;
    queue_reset(iq);
    for (;;) {
        uint64_t act = __ballot(busy);
        int nact = __popcll(act);
        if (!queue_empty(iq) && nact < threshold) {
            const uint32_t id = queue_take(iq, work, total, (uint32_t)(64 - nact), lanes_below(~act),
                                           !busy, lane);
            if (!busy) {
                if (id != ~0u) {
                    uint32_t lo = 0, hi = nbat; // last batch with prefix <= id
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (prefix[mid] <= id) lo = mid;
                        else hi = mid;
                    }
                    const vg_batch b = bat[lo];
                    const uint32_t j = id - prefix[lo];
                    if (j < b.n) { // ids past a batch's n (an inconsistent prefix) are ignored
                    const float4 *rp = reinterpret_cast<const float4 *>(b.rays + j);
                    const float4 a = rp[0], c = rp[1];
                    const float o[3] = {a.x, a.y, a.z}, din[3] = {c.x, c.y, c.z};
                    float d[3];
                    out = b.hits + j;
                    f = b.flags | __float_as_uint(c.w);
                    tmax = a.w;
                    lod = (f & VG_RAY_LOD) && !(f & (VG_RAY_SKIP_TRANSPARENT | VG_RAY_ANY_HIT));
                    hit_miss(h);
                    if (normalize_or_zero(din, d) && walk_init(r, w, o, d)) {
                        phase_reset(q);
                        it = 0;
                        busy = true;
                    } else {
                        store_hit(out, h);
                    }
                    }
                }
            }
            nact = __popcll(__ballot(busy));
        }
        if (nact == 0) {
            if (queue_empty(iq)) break;
            continue;
        }
        do {
            if (busy) {
                bool done = walk_iter<COUNT, FOLD>(r, q, w, tmax, f, lod, h, cnt);
                done = done || ++it >= max_iter;
                if (done) {
                    store_hit(out, h);
                    busy = false;
                }
            }
        } while (__popcll(__ballot(busy)) >= threshold);
    }
    if (COUNT) flush_counts(cnt, counts);
}

// Upload relayout: reference-order bricks (72 B AoS) -> nested occupancy + materials.
__global__ void k_relayout_bricks(const vg_brick *__restrict__ staged,
                                  const uint8_t *__restrict__ bu_ref, int64_t first, int64_t count,
                                  int32_t bx_n, int32_t by_n, int32_t cx_n, int32_t cy_n,
                                  uint64_t *__restrict__ occ, uint8_t *__restrict__ mat,
                                  uint8_t *__restrict__ bu) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int64_t bi = first + i;
    int32_t bx = (int32_t)(bi % bx_n);
    int32_t by = (int32_t)((bi / bx_n) % by_n);
    int32_t bz = (int32_t)(bi / ((int64_t)bx_n * by_n));
    int64_t ci = (bx >> 4) + (int64_t)(by >> 4) * cx_n + (int64_t)(bz >> 4) * cx_n * cy_n;
    int64_t tn = ci * 64 + cbit((bx >> 2) & 3, (by >> 2) & 3, (bz >> 2) & 3);
    int64_t bn = tn * 64 + cbit(bx & 3, by & 3, bz & 3);
    const uint64_t *src = reinterpret_cast<const uint64_t *>(staged + i);
    occ[bn] = src[0];
    uint64_t *dst = reinterpret_cast<uint64_t *>(mat + bn * 64);
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = src[1 + k];
    bu[bn] = bu_ref[i]; // the staging buffer holds this slab only
}

// Plain 16-byte-per-lane copy: the calibration kernel of known traffic (bytes read = bytes
// written = n16 * 16) that HBM counters of the other kernels are scaled by.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ void k_relayout_tiles(const uint64_t *__restrict__ tile_ref,
                                 const uint8_t *__restrict__ tu_ref, int64_t n, int32_t tx_n,
                                 int32_t ty_n, int32_t cx_n, int32_t cy_n,
                                 uint64_t *__restrict__ tile, uint8_t *__restrict__ tu) {
    int64_t ti = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ti >= n) return;
    int32_t tx = (int32_t)(ti % tx_n);
    int32_t ty = (int32_t)((ti / tx_n) % ty_n);
    int32_t tz = (int32_t)(ti / ((int64_t)tx_n * ty_n));
    int64_t ci = (tx >> 2) + (int64_t)(ty >> 2) * cx_n + (int64_t)(tz >> 2) * cx_n * cy_n;
    int64_t tn = ci * 64 + cbit(tx & 3, ty & 3, tz & 3);
    tile[tn] = tile_ref[ti];
    tu[tn] = tu_ref[ti];
}

} // namespace vg

using namespace vg;

namespace vg {

static int ensure_beam_capacity(vg_world *w, size_t blocks) {
    if (blocks <= w->beam_cap) return VG_OK;
    HIP_TRY(hipDeviceSynchronize());
    if (w->beam_t) HIP_TRY(hipFree(w->beam_t));
    w->beam_t = nullptr;
    w->beam_cap = 0;
    if (hipMalloc((void **)&w->beam_t, blocks * sizeof(float)) != hipSuccess)
        return fail(VG_E_NOMEM, "beam buffer of %zu blocks", blocks);
    w->beam_cap = blocks;
    return VG_OK;
}

// The persistent frame launch over caller-provided scratch: `work` (the ray counter, reset
// by the first kernel) and `beam_t` (one float per 8x8 block). Allocates nothing.
static int launch_frame_persistent_on(vg_world *w, const CamParams &cp, int32_t width,
                                      int32_t height, float ptmax, uint32_t pflags,
                                      const float s[3], float stmax, vg_hit *hits,
                                      uint8_t *occluded, bool shadow, bool beam, bool count,
                                      hipStream_t stream, uint32_t *work, float *beam_t,
                                      const uint32_t *list, const uint32_t *list_n) {
    // a block list names 8x8 blocks (vg_frame_dirty_blocks); else the context's block shape
    const uint32_t bws = list ? 3u : (uint32_t)w->ctx->frame_bws;
    const uint32_t bw = 1u << bws, bh = 64u >> bws;
    const uint32_t nbx = ((uint32_t)width + bw - 1) / bw, nby = ((uint32_t)height + bh - 1) / bh;
    const uint64_t nblocks = (uint64_t)nbx * nby;
    unsigned long long *cd = count ? w->counts_dev : nullptr;
    FrameParams fp;
    fp.cam = cp;
    for (int k = 0; k < 3; ++k) {
        fp.sun[k] = s[k];
        // the device's safe_inv, in host IEEE single precision (both correctly rounded)
        fp.sun_inv[k] = std::fabs(s[k]) < 1e-8f ? 1e30f : 1.0f / s[k];
    }
    fp.width = width;
    fp.height = height;
    fp.nbx = nbx;
    fp.bws = bws;
    const dim3 grid((uint32_t)persistent_grid(w->ctx, w->ctx->frame_wpc)), block(kBlock);
    const uint32_t stripes = (uint32_t)w->ctx->frame_stripes;
    fp.grid_wg = grid.x;
    fp.stripes = stripes;
    fp.nblk = list ? 0u : (uint32_t)nblocks;
    if (beam) {
        dim3 g((uint32_t)((nblocks + 3) / 4));
        if (count)
            hipLaunchKernelGGL((k_beam<true>), g, dim3(kBlock), 0, stream, w->dw, fp,
                               (uint32_t)nblocks, ptmax, beam_t, work, list, list_n, cd);
        else
            hipLaunchKernelGGL((k_beam<false>), g, dim3(kBlock), 0, stream, w->dw, fp,
                               (uint32_t)nblocks, ptmax, beam_t, work, list, list_n, cd);
    } else {
        hipLaunchKernelGGL(k_frame_setup, dim3(1), dim3(kMaxStripes), 0, stream, work, fp);
    }
    HIP_TRY(hipGetLastError());
    const uint32_t total = (uint32_t)(nblocks * 64);
    const int thr = w->ctx->frame_refill;
#define LAUNCH_P(C, SH, B)                                                                     \
    do {                                                                                       \
        if (w->dw.fold)                                                                        \
            hipLaunchKernelGGL((k_frame_persistent<C, SH, B, true>), grid, block, 0, stream,   \
                               w->dw, total, ptmax, pflags, stmax, beam_t, hits, occluded,     \
                               work, thr, stripes, list, list_n, cd);                          \
        else                                                                                   \
            hipLaunchKernelGGL((k_frame_persistent<C, SH, B, false>), grid, block, 0, stream,  \
                               w->dw, total, ptmax, pflags, stmax, beam_t, hits, occluded,     \
                               work, thr, stripes, list, list_n, cd);                          \
    } while (0)
    if (count) {
        if (shadow) {
            if (beam) LAUNCH_P(true, true, true);
            else LAUNCH_P(true, true, false);
        } else {
            if (beam) LAUNCH_P(true, false, true);
            else LAUNCH_P(true, false, false);
        }
    } else {
        if (shadow) {
            if (beam) LAUNCH_P(false, true, true);
            else LAUNCH_P(false, true, false);
        } else {
            if (beam) LAUNCH_P(false, false, true);
            else LAUNCH_P(false, false, false);
        }
    }
#undef LAUNCH_P
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

// Pixel blocks (one wave's 64 rays each) of a width x height frame: the context's block
// shape, or 8x8 for a block list (the dirty-block masks name 8x8 blocks).
static uint64_t frame_blocks(const vg_ctx *c, int32_t width, int32_t height, bool list) {
    const uint32_t bws = list ? 3u : (uint32_t)c->frame_bws;
    const uint64_t bw = 1u << bws, bh = 64u >> bws;
    return (((uint64_t)width + bw - 1) / bw) * (((uint64_t)height + bh - 1) / bh);
}

static int launch_frame_persistent(vg_world *w, const CamParams &cp, int32_t width, int32_t height,
                                   float ptmax, uint32_t pflags, const float s[3], float stmax,
                                   vg_hit *hits, uint8_t *occluded, bool shadow, bool beam,
                                   bool count, hipStream_t stream,
                                   const uint32_t *list = nullptr,
                                   const uint32_t *list_n = nullptr) {
    const uint64_t nblocks = frame_blocks(w->ctx, width, height, list != nullptr);
    // the counter may overshoot the ray count by up to 64 per wave after it runs out
    if (nblocks * 64 > 0xF0000000ull) return fail(VG_E_INVALID, "image too large");
    if (beam) {
        int st = ensure_beam_capacity(w, (size_t)nblocks);
        if (st) return st;
    }
    return launch_frame_persistent_on(w, cp, width, height, ptmax, pflags, s, stmax, hits,
                                      occluded, shadow, beam, count, stream, w->work, w->beam_t,
                                      list, list_n);
}

// Scratch of the no-allocation frame entry point (vg_trace_frame_noalloc): one array per
// device, `scratch`, holding the striped ray counters and the launch parameters (kWorkBytes)
// and from there the beam start t of up to kNoAllocBlocks 8x8 pixel blocks. Kernels get its
// device address, taken once per device with hipGetSymbolAddress (the name itself is a
// host-side stand-in). It keeps the last call's contents; the first kernel of every call
// resets the counters and writes the parameters.
constexpr size_t kNoAllocBlocks = 512 * 512; // 8x8 blocks of a 4096x4096 image
constexpr size_t kNoAllocBytes = kWorkBytes + kNoAllocBlocks * sizeof(float);
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

struct NoAllocScratch {
    uint32_t *work;
    float *beam_t;
};

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// The device address of `scratch` on the current device `device` (taken on the first call
// for that device, under a lock).
static int noalloc_scratch(int device, NoAllocScratch *out) {
    if (device < 0 || device >= kMaxDevices) return fail(VG_E_INVALID, "device %d", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(VG_E_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    uint8_t *base = static_cast<uint8_t *>(g_scratch_ptr[device]);
    out->work = reinterpret_cast<uint32_t *>(base);
    out->beam_t = reinterpret_cast<float *>(base + kWorkBytes);
    return VG_OK;
}

static int launch_frame(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                        float ptmax, uint32_t pflags, const float sun[3], float stmax,
                        vg_hit *hits, uint8_t *occluded, int mode, bool count,
                        hipStream_t stream, const NoAllocScratch *na = nullptr) {
    int st = check_world(w);
    if (st) return st;
    if (!camera || width <= 0 || height <= 0 || !hits) return fail(VG_E_INVALID, "bad frame args");
    if (camera->resolution[0] != (float)width || camera->resolution[1] != (float)height)
        return fail(VG_E_INVALID, "camera resolution %gx%g does not match %dx%d",
                    camera->resolution[0], camera->resolution[1], width, height);
    float s[3] = {0.0f, 0.0f, 0.0f};
    if (mode != 0) {
        if (!sun || !occluded) return fail(VG_E_INVALID, "null sun/occluded");
        if (!host_normalize(sun, s)) s[0] = s[1] = s[2] = 0.0f;
    }
    CamParams cp = cam_params(camera);
    unsigned long long *cd = count ? w->counts_dev : nullptr;
    const bool beam = (pflags & VG_RAY_BEAM) != 0;
    if (w->ctx->schedule == VG_SCHEDULE_PERSISTENT && mode != 2) {
        if (na) // the caller checked the block count against kNoAllocBlocks
            return launch_frame_persistent_on(w, cp, width, height, ptmax, pflags, s, stmax,
                                              hits, occluded, mode == 1, beam, count, stream,
                                              na->work, na->beam_t, nullptr, nullptr);
        return launch_frame_persistent(w, cp, width, height, ptmax, pflags, s, stmax, hits,
                                       occluded, mode == 1, beam, count, stream);
    }
    const uint32_t nbx = (uint32_t)((width + 15) / 16), nby = (uint32_t)((height + 15) / 16);
    const uint32_t nb = nbx * nby;
    dim3 grid(nb), block(kBlock);
#define LAUNCH_F(C, M, B)                                                                     \
    hipLaunchKernelGGL((k_trace_frame<C, M, B>), grid, block, 0, stream, w->dw, cp, width,      \
                       height, nbx, ptmax, pflags, s[0], s[1], s[2], stmax, hits, occluded, cd)
    if (count) {
        if (mode == 0) {
            if (beam) LAUNCH_F(true, 0, true);
            else LAUNCH_F(true, 0, false);
        } else if (mode == 1) {
            if (beam) LAUNCH_F(true, 1, true);
            else LAUNCH_F(true, 1, false);
        } else {
            LAUNCH_F(true, 2, false);
        }
    } else {
        if (mode == 0) {
            if (beam) LAUNCH_F(false, 0, true);
            else LAUNCH_F(false, 0, false);
        } else if (mode == 1) {
            if (beam) LAUNCH_F(false, 1, true);
            else LAUNCH_F(false, 1, false);
        } else {
            LAUNCH_F(false, 2, false);
        }
    }
#undef LAUNCH_F
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

static int launch_rays(vg_world *w, const vg_ray *rays, size_t n, uint32_t flags, vg_hit *hits,
                       bool count, hipStream_t stream) {
    int st = check_world(w);
    if (st) return st;
    if (n == 0) return VG_OK;
    if (!rays || !hits) return fail(VG_E_INVALID, "null rays/hits");
    if (((uintptr_t)rays & 15u) != 0) return fail(VG_E_INVALID, "rays must be 16-byte aligned");
    if (((uintptr_t)hits & 3u) != 0) return fail(VG_E_INVALID, "hits must be 4-byte aligned");
    // the counter may overshoot the ray count by one id chunk per wave after it runs out
    if (n > 0xF0000000ull) return fail(VG_E_INVALID, "too many rays");
    if (w->ctx->schedule == VG_SCHEDULE_PERSISTENT) {
        hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(1), 0, stream, w->work);
        const dim3 grid((uint32_t)persistent_grid(w->ctx, w->ctx->rays_wpc)), block(kBlock);
#define LAUNCH_R(C, F)                                                                         \
    hipLaunchKernelGGL((k_rays_persistent<C, F>), grid, block, 0, stream, w->dw, rays,           \
                       (uint32_t)n, flags, hits, w->work, w->ctx->rays_refill,                  \
                       C ? w->counts_dev : nullptr)
        if (count) {
            if (w->dw.fold) LAUNCH_R(true, true);
            else LAUNCH_R(true, false);
        } else {
            if (w->dw.fold) LAUNCH_R(false, true);
            else LAUNCH_R(false, false);
        }
#undef LAUNCH_R
        HIP_TRY(hipGetLastError());
        return VG_OK;
    }
    int64_t nb = ((int64_t)n + kBlock - 1) / kBlock;
    if (count)
        hipLaunchKernelGGL((k_trace_rays<true>), dim3((uint32_t)nb), dim3(kBlock), 0, stream,
                           w->dw, rays, (int64_t)n, flags, hits, w->counts_dev);
    else
        hipLaunchKernelGGL((k_trace_rays<false>), dim3((uint32_t)nb), dim3(kBlock), 0, stream,
                           w->dw, rays, (int64_t)n, flags, hits, nullptr);
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

// The y bound of an uploaded world (DevWorld::ytop): 4 x (the highest window-local brick row
// that holds a set tile bit, + 1), a uniform tile counting as all its rows; tiles in the
// reference order (x fastest, then y, then z). A window that folds y gets no bound.
static int32_t world_ytop(const Dims &d, const uint64_t *tile_mask, const uint8_t *tile_uniform,
                          int32_t ro_mod_y) {
    if (ro_mod_y != 0) return d.v[1];
    int32_t top = 0;
    for (int64_t t = 0; t < d.n_tiles; ++t) {
        const int32_t ty = (int32_t)((t / d.t[0]) % d.t[1]);
        if (tile_uniform[t]) top = std::max(top, ty * 16 + 16);
        else if (tile_mask[t])
            top = std::max(top, ty * 16 + ((63 - __builtin_clzll(tile_mask[t])) >> 4) * 4 + 4);
    }
    return top;
}

static void free_world(vg_world *w) {
    void *ptrs[] = {w->l4,     w->chunk,   w->tile,       w->occ,  w->tile_uni, w->brick_uni,
                    w->mat,    w->scratch, w->counts_dev, w->work, w->beam_t,   w->ytop_dev};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (w->wlist) (void)hipFree(w->wlist);
    if (w->wlist_n) (void)hipFree(w->wlist_n);
    vg_edit_state_free(w);
}

} // namespace vg

extern "C" {

int vg_ctx_create(int device, vg_ctx **out) {
    if (!out) return fail(VG_E_INVALID, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(VG_E_DEVICE, "no HIP device (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(VG_E_INVALID, "device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    vg_ctx *c = new (std::nothrow) vg_ctx();
    if (!c) return fail(VG_E_NOMEM, "ctx");
    c->device = device;
    c->cus = cus > 0 ? cus : 1;
    c->schedule = VG_SCHEDULE_PERSISTENT;
    c->frame_wpc = 16;
    c->frame_refill = 32;
    c->frame_stripes = 4;
    c->frame_bws = 3;
    c->rays_wpc = 16;
    c->rays_refill = 32;
    c->probe_wpc = 8;
    c->probe_refill = 8;
    c->probe_chunk = 0;
    c->probe_tbatch = 8;
    c->probe_matspec = 0;
    NoAllocScratch na; // the no-allocation entry point's scratch address on this device
    int st = noalloc_scratch(device, &na);
    if (st) {
        delete c;
        return st;
    }
    *out = c;
    return VG_OK;
}

int vg_ctx_destroy(vg_ctx *ctx) {
    delete ctx;
    return VG_OK;
}

int vg_world_create(vg_ctx *ctx, const int32_t dims[3], vg_world **out) {
    if (!ctx || !out) return fail(VG_E_INVALID, "null ctx/out");
    Dims d;
    int st = make_dims(dims, &d);
    if (st) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    vg_world *w = new (std::nothrow) vg_world();
    if (!w) return fail(VG_E_NOMEM, "world");
    std::memset(w, 0, sizeof(*w));
    w->ctx = ctx;
    w->d = d;
    struct {
        void **p;
        size_t bytes;
    } allocs[] = {
        {(void **)&w->l4, (size_t)d.n_l4 * 8},
        {(void **)&w->chunk, (size_t)d.n_chunks * 8},
        {(void **)&w->tile, (size_t)d.n_tiles * 8},
        {(void **)&w->tile_uni, (size_t)d.n_tiles},
        {(void **)&w->occ, (size_t)d.n_bricks * 8},
        {(void **)&w->brick_uni, (size_t)d.n_bricks},
        {(void **)&w->mat, (size_t)d.n_bricks * 64},
        {(void **)&w->scratch, 256},
        {(void **)&w->counts_dev, VG_IT_COUNT * sizeof(unsigned long long)},
        {(void **)&w->work, kWorkBytes},
        {(void **)&w->beam_t, (size_t)kBeamDefaultBlocks * sizeof(float)},
        {(void **)&w->ytop_dev, 256},
    };
    for (auto &a : allocs) {
        if (hipMalloc(a.p, a.bytes) != hipSuccess) {
            free_world(w);
            delete w;
            return fail(VG_E_NOMEM, "hipMalloc(%zu) failed", a.bytes);
        }
    }
    w->beam_cap = kBeamDefaultBlocks;
    DevWorld &dw = w->dw;
    dw.l4 = w->l4;
    dw.chunk = w->chunk;
    dw.tile = w->tile;
    dw.occ = w->occ;
    dw.tile_uni = w->tile_uni;
    dw.brick_uni = w->brick_uni;
    dw.mat = w->mat;
    for (int k = 0; k < 3; ++k) {
        dw.dims[k] = d.v[k];
        dw.cdim[k] = d.c[k];
        dw.l4dim[k] = d.l4[k];
        dw.origin[k] = 0;
        dw.ro_mod[k] = 0;
    }
    dw.fold = 0;
    // no bound until the world is uploaded or generated (a walk never ends early on it)
    dw.ytop = d.v[1];
    dw.ytop_p = w->ytop_dev;
    if (hipMemcpy(w->ytop_dev, &dw.ytop, sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
        free_world(w);
        delete w;
        return fail(VG_E_DEVICE, "world y bound");
    }
    *out = w;
    return VG_OK;
}

int vg_world_destroy(vg_world *w) {
    if (!w) return VG_OK;
    (void)hipSetDevice(w->ctx->device);
    free_world(w);
    delete w;
    return VG_OK;
}

int vg_world_upload(vg_world *w, const vg_brick *bricks, const uint64_t *tile_mask,
                    const uint64_t *chunk_mask, const uint64_t *l4_mask,
                    const uint8_t *brick_uniform, const uint8_t *tile_uniform,
                    const int32_t origin[3]) {
    if (!w) return fail(VG_E_INVALID, "null world");
    if (!bricks || !tile_mask || !chunk_mask || !l4_mask || !brick_uniform || !tile_uniform ||
        !origin)
        return fail(VG_E_INVALID, "null upload array");
    const Dims &d = w->d;
    for (int k = 0; k < 3; ++k)
        if (pos_mod(origin[k], d.v[k]) != 0 && d.v[k] % 256 != 0)
            return fail(VG_E_INVALID,
                        "origin[%d]=%d folds the window but dims[%d]=%d is not a multiple of 256", k,
                        origin[k], k, d.v[k]);
    int st = set_device(w);
    if (st) return st;
    w->uploaded = false;
    HIP_TRY(hipMemcpy(w->l4, l4_mask, (size_t)d.n_l4 * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(w->chunk, chunk_mask, (size_t)d.n_chunks * 8, hipMemcpyHostToDevice));
    // tiles
    {
        void *tmp = nullptr;
        size_t tbytes = (size_t)d.n_tiles * 9;
        HIP_TRY(hipMalloc(&tmp, tbytes));
        uint64_t *tref = (uint64_t *)tmp;
        uint8_t *turef = (uint8_t *)tmp + (size_t)d.n_tiles * 8;
        hipError_t e = hipMemcpy(tref, tile_mask, (size_t)d.n_tiles * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(turef, tile_uniform, (size_t)d.n_tiles, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            int64_t nb = (d.n_tiles + 255) / 256;
            hipLaunchKernelGGL(k_relayout_tiles, dim3((uint32_t)nb), dim3(256), 0, 0, tref, turef,
                               d.n_tiles, d.t[0], d.t[1], d.c[0], d.c[1], w->tile, w->tile_uni);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipDeviceSynchronize();
        }
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(VG_E_DEVICE, "tile upload: %s", hipGetErrorString(e));
    }
    // bricks, staged in slabs of at most 4 Mi bricks (288 MiB + 4 MiB)
    {
        const int64_t slab = std::min<int64_t>(d.n_bricks, (int64_t)1 << 22);
        void *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, (size_t)slab * (sizeof(vg_brick) + 1)));
        vg_brick *sb = (vg_brick *)tmp;
        uint8_t *sbu = (uint8_t *)tmp + (size_t)slab * sizeof(vg_brick);
        hipError_t e = hipSuccess;
        for (int64_t first = 0; first < d.n_bricks && e == hipSuccess; first += slab) {
            int64_t cnt = std::min(slab, d.n_bricks - first);
            e = hipMemcpy(sb, bricks + first, (size_t)cnt * sizeof(vg_brick), hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(sbu, brick_uniform + first, (size_t)cnt, hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                int64_t nb = (cnt + 255) / 256;
                hipLaunchKernelGGL(k_relayout_bricks, dim3((uint32_t)nb), dim3(256), 0, 0, sb, sbu,
                                   first, cnt, d.b[0], d.b[1], d.c[0], d.c[1], w->occ, w->mat,
                                   w->brick_uni);
                e = hipGetLastError();
                if (e == hipSuccess) e = hipDeviceSynchronize();
            }
        }
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(VG_E_DEVICE, "brick upload: %s", hipGetErrorString(e));
    }
    w->dw.fold = 0;
    for (int k = 0; k < 3; ++k) {
        w->dw.origin[k] = origin[k];
        w->dw.ro_mod[k] = pos_mod(origin[k], d.v[k]);
        if (w->dw.ro_mod[k]) w->dw.fold = 1;
    }
    const int32_t ytop = world_ytop(d, tile_mask, tile_uniform, w->dw.ro_mod[1]);
    HIP_TRY(hipMemcpy(w->ytop_dev, &ytop, sizeof(ytop), hipMemcpyHostToDevice));
    w->uploaded = true;
    return VG_OK;
}

int vg_trace_rays(vg_world *w, const vg_ray *rays, size_t n, uint32_t flags, vg_hit *hits,
                  void *stream) {
    return launch_rays(w, rays, n, flags, hits, false, (hipStream_t)stream);
}

int vg_trace_rays_counts(vg_world *w, const vg_ray *rays, size_t n, uint32_t flags, vg_hit *hits,
                         uint64_t counts_host[VG_IT_COUNT]) {
    if (!counts_host) return fail(VG_E_INVALID, "null counts");
    int st = check_world(w);
    if (st) return st;
    HIP_TRY(hipMemset(w->counts_dev, 0, VG_IT_COUNT * sizeof(unsigned long long)));
    st = launch_rays(w, rays, n, flags, hits, true, 0);
    if (st) return st;
    HIP_TRY(hipMemcpy(counts_host, w->counts_dev, VG_IT_COUNT * 8, hipMemcpyDeviceToHost));
    return VG_OK;
}

int vg_trace_primary(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                     float tmax, uint32_t flags, vg_hit *hits, void *stream) {
    return launch_frame(w, camera, width, height, tmax, flags, nullptr, 0.0f, hits, nullptr, 0,
                        false, (hipStream_t)stream);
}

int vg_trace_shadow(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                    const vg_hit *primary_hits, const float sun[3], float tmax, uint8_t *occluded,
                    void *stream) {
    return launch_frame(w, camera, width, height, 0.0f, 0u, sun, tmax,
                        const_cast<vg_hit *>(primary_hits), occluded, 2, false,
                        (hipStream_t)stream);
}

int vg_trace_frame(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                   float primary_tmax, uint32_t primary_flags, const float sun[3],
                   float shadow_tmax, vg_hit *hits, uint8_t *occluded, void *stream) {
    return launch_frame(w, camera, width, height, primary_tmax, primary_flags, sun, shadow_tmax,
                        hits, occluded, 1, false, (hipStream_t)stream);
}

int vg_trace_frame_noalloc(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                           float primary_tmax, uint32_t primary_flags, const float sun[3],
                           float shadow_tmax, vg_hit *hits, uint8_t *occluded, void *stream) {
    int st = check_world(w);
    if (st) return st;
    if (width <= 0 || height <= 0) return fail(VG_E_INVALID, "bad frame size %dx%d", width, height);
    const uint64_t nblocks = frame_blocks(w->ctx, width, height, false);
    if (nblocks > kNoAllocBlocks)
        return fail(VG_E_INVALID,
                    "%dx%d frame: %llu pixel blocks above the no-allocation scratch's %zu "
                    "(4096x4096 pixels); use vg_trace_frame",
                    width, height, (unsigned long long)nblocks, kNoAllocBlocks);
    NoAllocScratch na;
    st = noalloc_scratch(w->ctx->device, &na);
    if (st) return st;
    return launch_frame(w, camera, width, height, primary_tmax, primary_flags, sun, shadow_tmax,
                        hits, occluded, 1, false, (hipStream_t)stream, &na);
}

int vg_trace_water_secondary(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                             const vg_hit *primary_hits, vg_hit *refl_hits, vg_hit *refr_hits,
                             void *stream) {
    int st = check_world(w);
    if (st) return st;
    if (!camera || width <= 0 || height <= 0 || !primary_hits || !refl_hits || !refr_hits)
        return fail(VG_E_INVALID, "bad water-ray args");
    if (camera->resolution[0] != (float)width || camera->resolution[1] != (float)height)
        return fail(VG_E_INVALID, "camera resolution %gx%g does not match %dx%d",
                    camera->resolution[0], camera->resolution[1], width, height);
    if ((((uintptr_t)primary_hits | (uintptr_t)refl_hits | (uintptr_t)refr_hits) & 3u) != 0)
        return fail(VG_E_INVALID, "hit records must be 4-byte aligned");
    const uint32_t nbx = (uint32_t)((width + 7) / 8), nby = (uint32_t)((height + 7) / 8);
    const uint64_t nblocks = (uint64_t)nbx * nby;
    // the counter may overshoot the pixel count by up to 64 per wave after it runs out
    if (nblocks * 64 > 0xF0000000ull) return fail(VG_E_INVALID, "image too large");
    hipStream_t hs = (hipStream_t)stream;
    const size_t npix = (size_t)width * height;
    if (npix > w->wlist_cap) {
        HIP_TRY(hipDeviceSynchronize());
        if (w->wlist) (void)hipFree(w->wlist);
        w->wlist = nullptr;
        w->wlist_cap = 0;
        if (hipMalloc((void **)&w->wlist, npix * sizeof(uint32_t)) != hipSuccess)
            return fail(VG_E_NOMEM, "water pixel list");
        w->wlist_cap = npix;
    }
    if (!w->wlist_n && hipMalloc((void **)&w->wlist_n, 256) != hipSuccess)
        return fail(VG_E_NOMEM, "water pixel count");
    hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(1), 0, hs, w->wlist_n);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_water_compact, dim3((uint32_t)((nblocks * 64 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, hs, width, height, nbx, (uint32_t)(nblocks * 64),
                       primary_hits, refl_hits, refr_hits, w->wlist, w->wlist_n);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(1), 0, hs, w->work);
    HIP_TRY(hipGetLastError());
    const dim3 grid((uint32_t)persistent_grid(w->ctx, w->ctx->frame_wpc)), block(kBlock);
    hipLaunchKernelGGL(k_water_persistent, grid, block, 0, hs, w->dw, cam_params(camera), width,
                       w->wlist, w->wlist_n, primary_hits, refl_hits, refr_hits, w->work,
                       w->ctx->frame_refill);
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

int vg_trace_frame_blocks(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                          float primary_tmax, uint32_t primary_flags, const float sun[3],
                          float shadow_tmax, const uint32_t *mask, vg_hit *hits,
                          uint8_t *occluded, void *stream) {
    int st = check_world(w);
    if (st) return st;
    if (!camera || width <= 0 || height <= 0 || !hits || !mask)
        return fail(VG_E_INVALID, "bad frame args");
    if (camera->resolution[0] != (float)width || camera->resolution[1] != (float)height)
        return fail(VG_E_INVALID, "camera resolution %gx%g does not match %dx%d",
                    camera->resolution[0], camera->resolution[1], width, height);
    const bool shadow = sun != nullptr;
    float s[3] = {0.0f, 0.0f, 0.0f};
    if (shadow) {
        if (!occluded) return fail(VG_E_INVALID, "null occluded");
        if (!host_normalize(sun, s)) s[0] = s[1] = s[2] = 0.0f;
    }
    const uint32_t nbx = (uint32_t)((width + 7) / 8), nby = (uint32_t)((height + 7) / 8);
    const uint32_t *list = nullptr, *list_n = nullptr;
    st = edit_block_list(w, mask, nbx * nby, (hipStream_t)stream, &list, &list_n);
    if (st) return st;
    return launch_frame_persistent(w, cam_params(camera), width, height, primary_tmax,
                                   primary_flags, s, shadow_tmax, hits, occluded, shadow,
                                   (primary_flags & VG_RAY_BEAM) != 0, false, (hipStream_t)stream,
                                   list, list_n);
}

int vg_trace_frame_counts(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                          float primary_tmax, uint32_t primary_flags, const float sun[3],
                          float shadow_tmax, vg_hit *hits, uint8_t *occluded,
                          uint64_t counts_host[VG_IT_COUNT]) {
    if (!counts_host) return fail(VG_E_INVALID, "null counts");
    int st = check_world(w);
    if (st) return st;
    HIP_TRY(hipMemset(w->counts_dev, 0, VG_IT_COUNT * sizeof(unsigned long long)));
    st = launch_frame(w, camera, width, height, primary_tmax, primary_flags, sun, shadow_tmax, hits,
                      occluded, 1, true, 0);
    if (st) return st;
    HIP_TRY(hipMemcpy(counts_host, w->counts_dev, VG_IT_COUNT * 8, hipMemcpyDeviceToHost));
    return VG_OK;
}

int vg_raycast(vg_world *w, const float cam[3], const float dir[3], vg_hit *hit_host) {
    if (!cam || !dir || !hit_host) return fail(VG_E_INVALID, "null raycast argument");
    int st = check_world(w);
    if (st) return st;
    vg_ray r;
    std::memset(&r, 0, sizeof(r));
    for (int k = 0; k < 3; ++k) {
        r.o[k] = cam[k];
        r.d[k] = dir[k];
    }
    r.tmax = INFINITY; // the reference pick is bounded only by the window
    vg_ray *dr = (vg_ray *)w->scratch;
    vg_hit *dh = (vg_hit *)((char *)w->scratch + 64);
    HIP_TRY(hipMemcpy(dr, &r, sizeof(r), hipMemcpyHostToDevice));
    st = launch_rays(w, dr, 1, 0u, dh, false, 0);
    if (st) return st;
    HIP_TRY(hipMemcpy(hit_host, dh, sizeof(vg_hit), hipMemcpyDeviceToHost));
    return VG_OK;
}

int vg_device_count(int *out) {
    if (!out) return fail(VG_E_INVALID, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *out = n;
    return VG_OK;
}

#define CTX_DEVICE(ctx)                                                                       \
    do {                                                                                      \
        if (!(ctx)) return fail(VG_E_INVALID, "null ctx");                                    \
        HIP_TRY(hipSetDevice((ctx)->device));                                                 \
    } while (0)

int vg_alloc(vg_ctx *ctx, size_t bytes, void **dev_out) {
    CTX_DEVICE(ctx);
    if (!dev_out) return fail(VG_E_INVALID, "null out");
    *dev_out = nullptr;
    if (bytes == 0) return VG_OK;
    if (hipMalloc(dev_out, bytes) != hipSuccess) return fail(VG_E_NOMEM, "hipMalloc(%zu)", bytes);
    return VG_OK;
}

int vg_free(vg_ctx *ctx, void *dev) {
    CTX_DEVICE(ctx);
    if (dev) HIP_TRY(hipFree(dev));
    return VG_OK;
}

int vg_copy_to_device(vg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    CTX_DEVICE(ctx);
    if (bytes == 0) return VG_OK;
    if (!dst || !src) return fail(VG_E_INVALID, "null copy pointer");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return VG_OK;
}

int vg_copy_to_host(vg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    CTX_DEVICE(ctx);
    if (bytes == 0) return VG_OK;
    if (!dst || !src) return fail(VG_E_INVALID, "null copy pointer");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return VG_OK;
}

int vg_memset(vg_ctx *ctx, void *dev, int value, size_t bytes) {
    CTX_DEVICE(ctx);
    if (bytes == 0) return VG_OK;
    if (!dev) return fail(VG_E_INVALID, "null pointer");
    HIP_TRY(hipMemset(dev, value, bytes));
    return VG_OK;
}

int vg_copy_device(vg_ctx *ctx, void *dst, const void *src, size_t bytes) {
    CTX_DEVICE(ctx);
    if (bytes == 0) return VG_OK;
    if (!dst || !src) return fail(VG_E_INVALID, "null copy pointer");
    if (bytes % 16 == 0 && (uintptr_t)dst % 16 == 0 && (uintptr_t)src % 16 == 0) {
        const size_t n16 = bytes / 16;
        const size_t blocks = std::min<size_t>((n16 + 255) / 256, (size_t)ctx->cus * 16);
        hipLaunchKernelGGL(k_copy16, dim3((uint32_t)blocks), dim3(256), 0, 0,
                           (const uint4 *)src, (uint4 *)dst, n16);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
        return VG_OK;
    }
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return VG_OK;
}

int vg_synchronize(vg_ctx *ctx) {
    CTX_DEVICE(ctx);
    HIP_TRY(hipDeviceSynchronize());
    return VG_OK;
}

int vg_stream_create(vg_ctx *ctx, void **out) {
    CTX_DEVICE(ctx);
    if (!out) return fail(VG_E_INVALID, "null out");
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = (void *)s;
    return VG_OK;
}

int vg_stream_destroy(vg_ctx *ctx, void *stream) {
    CTX_DEVICE(ctx);
    if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return VG_OK;
}

int vg_stream_synchronize(vg_ctx *ctx, void *stream) {
    CTX_DEVICE(ctx);
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return VG_OK;
}

int vg_event_create(vg_ctx *ctx, void **out) {
    CTX_DEVICE(ctx);
    if (!out) return fail(VG_E_INVALID, "null out");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *out = (void *)e;
    return VG_OK;
}

int vg_event_destroy(vg_ctx *ctx, void *event) {
    CTX_DEVICE(ctx);
    if (event) HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return VG_OK;
}

int vg_event_record(vg_ctx *ctx, void *event, void *stream) {
    CTX_DEVICE(ctx);
    if (!event) return fail(VG_E_INVALID, "null event");
    HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return VG_OK;
}

int vg_event_elapsed_ms(vg_ctx *ctx, void *start, void *end, float *ms) {
    CTX_DEVICE(ctx);
    if (!start || !end || !ms) return fail(VG_E_INVALID, "null event/out");
    HIP_TRY(hipEventSynchronize((hipEvent_t)end));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return VG_OK;
}

} // extern "C"

namespace vg {
// Time `launch` (already bound to its arguments) over every (waves per CU, refill) pair and
// keep the fastest in the context. Each candidate runs once to warm up, then `reps` times
// between events on one stream; its score is the fastest repetition.
} // namespace vg

extern "C" {

int vg_ctx_set_option(vg_ctx *ctx, int option, int value) {
    if (!ctx) return fail(VG_E_INVALID, "null ctx");
    switch (option) {
    case VG_OPT_SCHEDULE:
        if (value != VG_SCHEDULE_PERSISTENT && value != VG_SCHEDULE_STATIC)
            return fail(VG_E_INVALID, "schedule %d", value);
        ctx->schedule = value;
        return VG_OK;
    case VG_OPT_FRAME_WAVES_PER_CU:
    case VG_OPT_RAYS_WAVES_PER_CU:
    case VG_OPT_PROBE_WAVES_PER_CU:
        if (value < 1 || value > 64) return fail(VG_E_INVALID, "waves per CU %d not in 1..64", value);
        (option == VG_OPT_FRAME_WAVES_PER_CU ? ctx->frame_wpc
         : option == VG_OPT_RAYS_WAVES_PER_CU ? ctx->rays_wpc
                                              : ctx->probe_wpc) = value;
        return VG_OK;
    case VG_OPT_FRAME_REFILL:
    case VG_OPT_RAYS_REFILL:
    case VG_OPT_PROBE_REFILL:
        if (value < 1 || value > 64) return fail(VG_E_INVALID, "refill threshold %d not in 1..64", value);
        (option == VG_OPT_FRAME_REFILL ? ctx->frame_refill
         : option == VG_OPT_RAYS_REFILL ? ctx->rays_refill
                                        : ctx->probe_refill) = value;
        return VG_OK;
    case VG_OPT_FRAME_BLOCK_W:
        if (value != 4 && value != 8 && value != 16 && value != 32)
            return fail(VG_E_INVALID, "frame block width %d: 4, 8, 16 or 32", value);
        ctx->frame_bws = value == 4 ? 2 : value == 8 ? 3 : value == 16 ? 4 : 5;
        return VG_OK;
    case VG_OPT_FRAME_STRIPES:
        if (value < 1 || value > (int)kMaxStripes)
            return fail(VG_E_INVALID, "frame stripes %d not in 1..%u", value, kMaxStripes);
        ctx->frame_stripes = value;
        return VG_OK;
    case VG_OPT_PROBE_CHUNK:
        if (value != 0 && (value < 16 || value > 1024))
            return fail(VG_E_INVALID, "probe id chunk %d: 0 or 16..1024", value);
        ctx->probe_chunk = value;
        return VG_OK;
    case VG_OPT_PROBE_TBATCH:
        if (value < 1 || value > 64)
            return fail(VG_E_INVALID, "probe transition batch %d not in 1..64", value);
        ctx->probe_tbatch = value;
        return VG_OK;
    case VG_OPT_PROBE_MATSPEC:
        if (value != 0 && value != 1)
            return fail(VG_E_INVALID, "probe material prefetch %d: 0 or 1", value);
        ctx->probe_matspec = value;
        return VG_OK;
    default:
        return fail(VG_E_INVALID, "option %d cannot be set", option);
    }
}

int vg_ctx_get_option(vg_ctx *ctx, int option, int *value) {
    if (!ctx || !value) return fail(VG_E_INVALID, "null ctx/value");
    switch (option) {
    case VG_OPT_SCHEDULE: *value = ctx->schedule; return VG_OK;
    case VG_OPT_FRAME_WAVES_PER_CU: *value = ctx->frame_wpc; return VG_OK;
    case VG_OPT_FRAME_REFILL: *value = ctx->frame_refill; return VG_OK;
    case VG_OPT_FRAME_STRIPES: *value = ctx->frame_stripes; return VG_OK;
    case VG_OPT_FRAME_BLOCK_W: *value = 1 << ctx->frame_bws; return VG_OK;
    case VG_OPT_RAYS_WAVES_PER_CU: *value = ctx->rays_wpc; return VG_OK;
    case VG_OPT_RAYS_REFILL: *value = ctx->rays_refill; return VG_OK;
    case VG_OPT_PROBE_WAVES_PER_CU: *value = ctx->probe_wpc; return VG_OK;
    case VG_OPT_PROBE_REFILL: *value = ctx->probe_refill; return VG_OK;
    case VG_OPT_PROBE_CHUNK: *value = ctx->probe_chunk; return VG_OK;
    case VG_OPT_PROBE_TBATCH: *value = ctx->probe_tbatch; return VG_OK;
    case VG_OPT_PROBE_MATSPEC: *value = ctx->probe_matspec; return VG_OK;
    case VG_OPT_CU_COUNT: *value = ctx->cus; return VG_OK;
    default: return fail(VG_E_INVALID, "unknown option %d", option);
    }
}

int vg_autotune_frame(vg_world *w, const vg_camera *camera, int32_t width, int32_t height,
                      float primary_tmax, uint32_t primary_flags, const float sun[3],
                      float shadow_tmax, vg_hit *hits, uint8_t *occluded, float *best_ms) {
    int st = check_world(w);
    if (st) return st;
    vg_ctx *c = w->ctx;
    // the frame kernel: pixel-block shape x waves per CU x refill threshold x work stripes
    // (the coherent frame refills best from a few lanes, measured; wider refill thresholds
    // only lose)
    static const int kWpc[] = {12, 16, 20, 24};
    static const int kRefill[] = {4, 8, 16};
    static const int kStripes[] = {1, 2, 4, 8};
    static const int kBws[] = {3, 2, 4, 5}; // block widths 8, 4, 16, 32
    float best = INFINITY;
    const int ow = c->frame_wpc, orf = c->frame_refill, os = c->frame_stripes, ob = c->frame_bws;
    int bw = ow, br = orf, bs = os, bb = ob;
    for (int b : kBws) {
        c->frame_bws = b;
        if (frame_blocks(c, width, height, false) * 64 > 0xF0000000ull) continue;
        float ms = INFINITY;
        int st = autotune_lists(c, &c->frame_wpc, TuneList{kWpc, 4}, &c->frame_refill,
                                TuneList{kRefill, 3}, &c->frame_stripes, TuneList{kStripes, 4},
                                [&](hipStream_t s) {
                                    return vg_trace_frame(w, camera, width, height,
                                                          primary_tmax, primary_flags, sun,
                                                          shadow_tmax, hits, occluded,
                                                          (void *)s);
                                },
                                &ms);
        if (st) { // leave the options as they were
            c->frame_wpc = ow;
            c->frame_refill = orf;
            c->frame_stripes = os;
            c->frame_bws = ob;
            return st;
        }
        if (ms < best) {
            best = ms;
            bw = c->frame_wpc;
            br = c->frame_refill;
            bs = c->frame_stripes;
            bb = b;
        }
    }
    c->frame_wpc = bw;
    c->frame_refill = br;
    c->frame_stripes = bs;
    c->frame_bws = bb;
    if (best_ms) *best_ms = best;
    return VG_OK;
}

int vg_autotune_rays(vg_world *w, const vg_ray *rays, size_t n, uint32_t flags, vg_hit *hits,
                     float *best_ms) {
    int st = check_world(w);
    if (st) return st;
    vg_ctx *c = w->ctx;
    return autotune(c, &c->rays_wpc, &c->rays_refill, [&](hipStream_t s) {
        return vg_trace_rays(w, rays, n, flags, hits, (void *)s);
    }, best_ms);
}

int vg_trace_batches(vg_world *w, const vg_batch *batches, const uint32_t *prefix,
                     uint32_t n_batches, uint32_t total, void *stream) {
    int st = check_world(w);
    if (st) return st;
    if (total == 0 || n_batches == 0) return VG_OK;
    if (!batches || !prefix) return fail(VG_E_INVALID, "null batches/prefix");
    if (total > 0xF0000000u) return fail(VG_E_INVALID, "too many rays");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(1), 0, s, w->work);
    const dim3 grid((uint32_t)persistent_grid(w->ctx, w->ctx->rays_wpc)), block(kBlock);
    if (w->dw.fold)
        hipLaunchKernelGGL((k_batches_persistent<false, true>), grid, block, 0, s, w->dw, batches,
                           prefix, n_batches, total, w->work, w->ctx->rays_refill, nullptr);
    else
        hipLaunchKernelGGL((k_batches_persistent<false, false>), grid, block, 0, s, w->dw, batches,
                           prefix, n_batches, total, w->work, w->ctx->rays_refill, nullptr);
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

int vg_trace_batches_counts(vg_world *w, const vg_batch *batches, const uint32_t *prefix,
                            uint32_t n_batches, uint32_t total, uint64_t counts_host[VG_IT_COUNT]) {
    if (!counts_host) return fail(VG_E_INVALID, "null counts");
    int st = check_world(w);
    if (st) return st;
    HIP_TRY(hipMemset(w->counts_dev, 0, VG_IT_COUNT * sizeof(unsigned long long)));
    if (total && n_batches) {
        if (!batches || !prefix) return fail(VG_E_INVALID, "null batches/prefix");
        hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(1), 0, 0, w->work);
        const dim3 grid((uint32_t)persistent_grid(w->ctx, w->ctx->rays_wpc)), block(kBlock);
        if (w->dw.fold)
            hipLaunchKernelGGL((k_batches_persistent<true, true>), grid, block, 0, 0, w->dw,
                               batches, prefix, n_batches, total, w->work, w->ctx->rays_refill,
                               w->counts_dev);
        else
            hipLaunchKernelGGL((k_batches_persistent<true, false>), grid, block, 0, 0, w->dw,
                               batches, prefix, n_batches, total, w->work, w->ctx->rays_refill,
                               w->counts_dev);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpy(counts_host, w->counts_dev, VG_IT_COUNT * 8, hipMemcpyDeviceToHost));
    return VG_OK;
}

int vg_autotune_batches(vg_world *w, const vg_batch *batches, const uint32_t *prefix,
                        uint32_t n_batches, uint32_t total, float *best_ms) {
    int st = check_world(w);
    if (st) return st;
    vg_ctx *c = w->ctx;
    return autotune(c, &c->rays_wpc, &c->rays_refill, [&](hipStream_t s) {
        return vg_trace_batches(w, batches, prefix, n_batches, total, (void *)s);
    }, best_ms);
}

struct vg_graph {
    int device;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int vg_graph_begin(vg_ctx *ctx, void *stream) {
    CTX_DEVICE(ctx);
    if (!stream) return fail(VG_E_INVALID, "capture needs a stream (not the null stream)");
    HIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return VG_OK;
}

int vg_graph_end(vg_ctx *ctx, void *stream, vg_graph **out) {
    CTX_DEVICE(ctx);
    if (!stream || !out) return fail(VG_E_INVALID, "null stream/out");
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(VG_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    vg_graph *gr = new (std::nothrow) vg_graph;
    if (!gr) {
        (void)hipGraphExecDestroy(ex);
        (void)hipGraphDestroy(g);
        return fail(VG_E_NOMEM, "graph");
    }
    gr->device = ctx->device;
    gr->graph = g;
    gr->exec = ex;
    *out = gr;
    return VG_OK;
}

int vg_graph_launch(vg_graph *g, void *stream) {
    if (!g) return fail(VG_E_INVALID, "null graph");
    HIP_TRY(hipSetDevice(g->device));
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return VG_OK;
}

int vg_graph_destroy(vg_graph *g) {
    if (!g) return VG_OK;
    (void)hipSetDevice(g->device);
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
    return VG_OK;
}

} // extern "C"
