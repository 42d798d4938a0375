// GI probe update for gfx950: cs_gi_probe_update (shaders/gi_probes.wgsl:274-396) as
// three launches on one stream.
//   k_probe_prep    one thread per updated probe: its world grid cell, whether the slot is
//                   new, its (relocated) ray origin and whether it is buried in a solid voxel.
//   k_probe_paths   persistent, ray-compacting: one lane per gather sample runs the whole
//                   path (gather ray -> sun shadow ray at the hit -> optional cosine bounce
//                   ray -> sun shadow ray at the bounce hit) through the exact walk, refilling
//                   idle lanes like the frame kernel, and writes the sample's gather face
//                   and, for a hit, its radiance and visible distance.
//   k_probe_reduce  a group of lanes per updated probe: the sky radiance of samples without
//                   a gather hit, SH-L1 projection of its samples in a fixed lane order
//                   (deterministic), the nearest hit for relocation, the staging sum and the
//                   hysteresis fold of complete cycles.
// Samples of one probe have consecutive ids, so a wave's gather rays share their origin.
// The learned update (vg_probes_update_learned) runs the path kernel without the bounce:
// every gather hit is appended to a compacted list of cache queries, one inference launch
// of the radiance cache adds its bounce tail. The same path launch also traces the training
// targets (cosine bounce + sun shadow) of a batch drawn from the previous learned update's
// hits (the list is double-buffered; the half in use is a device-side parity), so the long
// training paths overlap the gather work instead of forming a launch of their own.
#include "vg_nrc.h"

struct vg_probes {
    vg_world *world;
    int device;             // the world's device (destroy does not touch the world)
    int32_t grid[3];
    int32_t rays, epochs;
    int64_t total;          // probes
    vg_probe *probes;       // total records
    float4 *dirs;           // rays * epochs gather directions, then as many safe_inv entries
    float4 *bounce;         // VG_GI_BOUNCE_TABLE cosine-hemisphere directions, then their
                            // safe_inv entries
    float4 *palette;        // 256 albedo entries
    float4 *pwork;          // per updated probe: ray origin xyz, w = flags bits
    float4 *rad;            // per sample with a gather hit: radiance rgb, visible distance
    uint8_t *face;          // per sample: face code of the gather hit (VG_FACE_NONE: none)
    vg_nrc_query *queries;  // learned update: compacted gather hits, two halves of `ns` entries
    size_t ns;              // samples per update at most (total probes x rays)
    uint32_t *ctr;          // device counters, see kGiCtrs
    // Heavy paths (walk iterations >= a threshold) of the last update, traced first by the
    // next one, which skips them in the natural order: two halves, alternating per update
    uint32_t *heavy;        // sample ids: 2 * hcap
    uint32_t *hmask;        // one bit per sample: 2 * hwords
    uint32_t hcap, hwords;
    // per-launch timing (vg_probes_set_timing): 5 events per recorded update, bracketing prep,
    // the path kernel, the cache inference and the reduction on the update's stream
    hipEvent_t *tev;        // kTimingCap * 5 events (null: timing off)
    int tn;                 // updates recorded
    unsigned long long *counts; // VG_IT_COUNT walk classes + 3 ray kinds
};

namespace vg {

constexpr uint32_t kProbeNew = 1u, kProbeBuried = 2u;
constexpr int kGiCounters = VG_IT_COUNT + 3;
// Device counters of a probe grid: [0] path work, [1] this update's cache queries, [2] the
// previous learned update's query count, [3] the query-list half this update writes,
// [4 + s - 1] work of training step s >= 1 (vg_nrc_config.steps <= 64), [kCtrHeavy] this
// update's heavy paths, [kCtrHeavy + 1] the previous update's, [kCtrUpd] the number of updates
// so far (its parity picks the heavy-list halves; the reduction, the update's last launch,
// counts it up, so no host-side state changes per update and a graph may capture one).
constexpr int kCtrHeavy = 4 + 64;
constexpr int kCtrUpd = kCtrHeavy + 2;
constexpr int kTimingCap = 64; // updates one timing record holds
constexpr int kGiCtrs = kCtrUpd + 1;
// Striped sample queue of the path kernel: kGiStripes counters, one 128-byte line each, after
// the counters above. Stripe s holds a contiguous range of whole probes (probe slots run x
// fastest, then y, then z, so a stripe is a slab of the grid); workgroups go round-robin over
// the 8 XCDs, so the waves of one XCD start on one stripe and keep its slab of the world in
// their L2.
constexpr int kGiStripes = 8;
constexpr int kGiStripe0 = 96;
constexpr int kGiStripeWords = 32;
constexpr int kGiCtrWords = kGiStripe0 + kGiStripes * kGiStripeWords;
static_assert(kGiCtrs <= kGiStripe0, "counters overlap the stripes");

// Static first reservations of the probe path launch (its waves 4 per workgroup, workgroup b
// on stripe b % kGiStripes): wave gw starts with the ids [64 gw, 64 gw + 64) of the front
// region (training samples, then the previous update's heavy samples) below lim = min(front,
// 64 waves); each wave past those (gw >= nfw = ceil(lim / 64)) starts with the next 64 natural
// samples of its stripe, in the order of gw. k_probe_prep starts the shared counters past the
// static ids, so no wave waits on an atomic for its first 64 samples (at the launch's start
// thousands of waves on one counter take ~20 us, and hold up the memory requests behind them).
// The waves of stripe sc with an index below x:
__device__ __forceinline__ uint32_t stripe_waves_below(uint32_t x, uint32_t sc) {
    const uint32_t b = x >> 2, rem = x & 3u; // whole workgroups below x, waves of the next one
    const uint32_t full = b > sc ? (b - sc + kGiStripes - 1u) / kGiStripes : 0u;
    return full * 4u + (b % kGiStripes == sc ? rem : 0u);
}

// Natural samples [s0, s1) of stripe sc (whole probes, slen samples per stripe).
__device__ __forceinline__ void stripe_range(uint32_t sc, uint32_t slen, uint32_t n, uint32_t &s0,
                                             uint32_t &s1) {
    s0 = sc * slen < n ? sc * slen : n;
    s1 = n - s0 < slen ? n : s0 + slen;
}

// ----------------------------------------------------------------- helpers --
__device__ __forceinline__ float gi_smooth(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

// sky_color (shaders/raymarch.wgsl:2532-2579) for a daytime or dusk sun: horizon-zenith
// blend, the dusk warm wash, the forward-scatter halo and the sun disc. The night stars
// term is not evaluated.
__device__ __forceinline__ void sky_color(const vg_gi_params &p, const float d[3], float out[3]) {
    const float up = d[1] < -0.2f ? -0.2f : (d[1] > 1.0f ? 1.0f : d[1]);
    const float zf = gi_smooth(0.0f, 0.55f, up);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = p.sky_horizon[k] * (1.0f - zf) + p.sky_zenith[k] * zf;
    if (p.dusk_w > 0.0f) {
        const float l = sqrtf(d[0] * d[0] + d[2] * d[2]);
        float azi = 0.0f;
        if (l > 0.0f) {
            azi = (d[0] / l) * p.sun_xz[0] + (d[2] / l) * p.sun_xz[1];
            azi = azi < 0.0f ? 0.0f : (azi > 1.0f ? 1.0f : azi);
        }
        const float wgt = p.dusk_w * (azi * azi * azi) * (1.0f - zf);
        out[0] += 1.50f * wgt;
        out[1] += 0.55f * wgt;
        out[2] += 0.18f * wgt;
    }
    float sa = d[0] * p.sun[0] + d[1] * p.sun[1] + d[2] * p.sun[2];
    sa = sa > 0.0f ? sa : 0.0f;
    const float s2 = sa * sa;
    const float h = s2 * s2 * s2 * p.halo;
    out[0] += 1.0f * h;
    out[1] += 0.78f * h;
    out[2] += 0.52f * h;
    if (sa > 0.9985f) {
        out[0] = out[0] * 0.1f + 2.4f * 0.9f;
        out[1] = out[1] * 0.1f + 2.0f * 0.9f;
        out[2] = out[2] * 0.1f + 1.5f * 0.9f;
    }
}

// Outward normal component of face code f on axis k.
__device__ __forceinline__ float face_n(uint32_t f, int k) {
    return (f < 6u && (int)(f >> 1) == k) ? ((f & 1u) ? -1.0f : 1.0f) : 0.0f;
}

// max(0, n . sun) for an axis-aligned face normal.
__device__ __forceinline__ float face_ndl(uint32_t f, const float s[3]) {
    if (f >= 6u) return 0.0f;
    const float v = (f & 1u) ? -s[f >> 1] : s[f >> 1];
    return v > 0.0f ? v : 0.0f;
}

// Secondary ray origin off a hit: (o + d t) + n * 0.03 (shaders/gi_probes.wgsl:239).
__device__ __forceinline__ void gi_offset_origin(const float o[3], const float d[3], float t,
                                                 uint32_t f, float out[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = (o[k] + d[k] * t) + face_n(f, k) * 0.03f;
}

// lowbias32 integer mixer (C. Wellons): the bounce-table index of a sample.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint32_t bounce_index(uint32_t seed, int32_t round, uint32_t probe,
                                                 uint32_t k) {
    uint32_t x = mix32(seed ^ ((uint32_t)round * 0x85EBCA6Bu));
    x = mix32(x ^ probe);
    x = mix32(x ^ k);
    return x >> 20; // VG_GI_BOUNCE_TABLE = 4096 entries
}

// Training sample j of a batch of n drawn from a list of nq hits: stratified over the list
// (whose order follows probe order) from a per-step random offset, and its bounce-table
// index.
__device__ __forceinline__ void train_pick(uint32_t tseed, int32_t round, uint32_t j, uint32_t n,
                                           uint32_t nq, uint32_t &qi, uint32_t &bidx) {
    const uint32_t x = mix32(tseed ^ ((uint32_t)round * 0x85EBCA6Bu));
    qi = (uint32_t)(((uint64_t)(x % nq) + (uint64_t)j * nq / n) % nq);
    bidx = mix32(x ^ mix32(j ^ 0x68E31DA4u)) >> 20;
}

// Cosine-hemisphere table entry (t, b, c) turned to the hit face: c along the outward
// normal, t and b on the next two axes in cyclic order.
__device__ __forceinline__ void bounce_dir(const float4 e, uint32_t f, float out[3]) {
    const int ax = (int)(f >> 1);
    const float sg = (f & 1u) ? -1.0f : 1.0f;
    out[ax] = sg * e.z;
    out[ax == 2 ? 0 : ax + 1] = e.x;
    out[ax == 0 ? 2 : ax - 1] = e.y;
}

// safe_inv of each component of bounce_dir(e, f): the same permutation and sign applied to
// the entry's inverses ei, except that a tiny normal component inverts to +1e30 whatever
// its sign (safe_inv's cut-off), so it keeps that value.
__device__ __forceinline__ void bounce_inv(const float4 e, const float4 ei, uint32_t f,
                                           float out[3]) {
    const int ax = (int)(f >> 1);
    const float sg = (f & 1u) ? -1.0f : 1.0f;
    out[ax] = fabsf(e.z) < 1e-8f ? 1e30f : sg * ei.z;
    out[ax == 2 ? 0 : ax + 1] = ei.x;
    out[ax == 0 ? 2 : ax - 1] = ei.y;
}

// Material of a world voxel through the pyramid (0 outside the window). Every level is a
// dense array indexed by the voxel's coordinates, so the seven words are loaded together (one
// round trip) and the test chain then runs on registers.
__device__ __forceinline__ uint32_t material_at(const DevWorld &w, const int32_t v[3]) {
    int32_t sl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int32_t l = v[k] - w.origin[k];
        if ((uint32_t)l >= (uint32_t)w.dims[k]) return 0u;
        int32_t s = l + w.ro_mod[k];
        sl[k] = s >= w.dims[k] ? s - w.dims[k] : s;
    }
    const uint32_t l4i = (uint32_t)(sl[0] >> 8) + (uint32_t)(sl[1] >> 8) * w.l4dim[0] +
                         (uint32_t)(sl[2] >> 8) * w.l4dim[0] * w.l4dim[1];
    const uint32_t ci = (uint32_t)(sl[0] >> 6) + (uint32_t)(sl[1] >> 6) * w.cdim[0] +
                        (uint32_t)(sl[2] >> 6) * w.cdim[0] * w.cdim[1];
    const int cb = cbit((sl[0] >> 6) & 3, (sl[1] >> 6) & 3, (sl[2] >> 6) & 3);
    const int tb = cbit((sl[0] >> 4) & 3, (sl[1] >> 4) & 3, (sl[2] >> 4) & 3);
    const int bb = cbit((sl[0] >> 2) & 3, (sl[1] >> 2) & 3, (sl[2] >> 2) & 3);
    const int vb = cbit(sl[0] & 3, sl[1] & 3, sl[2] & 3);
    const uint32_t tn = ci * 64u + (uint32_t)tb;
    const uint32_t bn = tn * 64u + (uint32_t)bb;
    const uint64_t l4w = w.l4[l4i], cw = w.chunk[ci], tw = w.tile[tn], ow = w.occ[bn];
    const uint32_t tu = w.tile_uni[tn], bu = w.brick_uni[bn];
    const uint32_t m = w.mat[(uint64_t)bn * 64u + (uint32_t)vb];
    if (!((l4w >> cb) & 1) || !((cw >> tb) & 1)) return 0u;
    if (tu) return tu;
    if (!((tw >> bb) & 1)) return 0u;
    if (bu) return bu;
    return ((ow >> vb) & 1) ? m : 0u;
}

// World grid cell of probe slot idx for the current window (shaders/gi_probes.wgsl:281-297).
__device__ __forceinline__ void probe_cell(const vg_gi_params &p, const DevWorld &w, int64_t idx,
                                           int32_t g[3]) {
    const int32_t sx = (int32_t)(idx % p.grid[0]);
    const int32_t sy = (int32_t)((idx / p.grid[0]) % p.grid[1]);
    const int32_t sz = (int32_t)(idx / ((int64_t)p.grid[0] * p.grid[1]));
    const int32_t ox = w.origin[0] / p.spacing, oz = w.origin[2] / p.spacing;
    int32_t mx = (sx - ox) % p.grid[0];
    mx = mx < 0 ? mx + p.grid[0] : mx;
    int32_t mz = (sz - oz) % p.grid[2];
    mz = mz < 0 ? mz + p.grid[2] : mz;
    g[0] = ox + mx;
    g[1] = sy;
    g[2] = oz + mz;
}

// ----------------------------------------------------------------- kernels --
__global__ __launch_bounds__(kBlock) void k_probe_prep(DevWorld w, vg_gi_params p, uint32_t n_upd,
                                                      uint32_t phase, const vg_probe *probes,
                                                      float4 *pwork, uint32_t *ctr, int learned,
                                                      uint32_t *hmask, uint32_t hwords,
                                                      uint32_t train_n, uint32_t hcap,
                                                      uint32_t static_ids) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j == 0) { // the kernels enqueued next consume the counters
        const uint32_t prev = ctr[1], par = ctr[3], hprev = ctr[kCtrHeavy];
        for (int k = 0; k < kCtrUpd; ++k) ctr[k] = 0u;
        // the shared counters start past the path launch's static reservations
        const uint32_t front = train_n + (hprev < hcap ? hprev : hcap);
        const uint32_t lim = front < static_ids ? front : static_ids;
        ctr[0] = lim;
        const uint32_t nw = static_ids / 64u, nfw = (lim + 63u) / 64u;
        const uint32_t N = (uint32_t)p.rays_per_probe, n = n_upd * N;
        const uint32_t slen = ((n_upd + kGiStripes - 1u) / kGiStripes) * N;
        for (uint32_t sc = 0; sc < (uint32_t)kGiStripes; ++sc) {
            uint32_t s0, s1;
            stripe_range(sc, slen, n, s0, s1);
            const uint32_t taken =
                (stripe_waves_below(nw, sc) - stripe_waves_below(nfw, sc)) * 64u;
            ctr[kGiStripe0 + sc * kGiStripeWords] = taken < s1 - s0 ? taken : s1 - s0;
        }
        ctr[2] = learned ? prev : 0u;
        ctr[3] = learned ? (par ^ 1u) : par;
        ctr[kCtrHeavy + 1] = hprev;
    }
    // this update's heavy-path bits start clear
    uint32_t *hmask_cur = hmask + (size_t)(ctr[kCtrUpd] & 1u) * hwords;
    for (uint32_t k = j; k < hwords; k += gridDim.x * kBlock) hmask_cur[k] = 0u;
    if (j >= n_upd) return;
    const int64_t idx = (int64_t)j * p.update_div + phase;
    int32_t g[3];
    probe_cell(p, w, idx, g);
    const float4 *rec = reinterpret_cast<const float4 *>(probes + idx);
    const float4 info = rec[5], st_info = rec[11];
    const bool is_new = info.w < 0.5f || info.x != (float)g[0] || info.y != (float)g[1] ||
                        info.z != (float)g[2];
    const float off[3] = {is_new ? 0.0f : st_info.y, is_new ? 0.0f : st_info.z,
                          is_new ? 0.0f : st_info.w};
    float o[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = ((float)g[k] + 0.5f) * (float)p.spacing + off[k];
    uint32_t fl = is_new ? kProbeNew : 0u;
    if (p.flags & VG_GI_SKIP_BURIED) {
        const int32_t v[3] = {(int32_t)floorf(o[0]), (int32_t)floorf(o[1]), (int32_t)floorf(o[2])};
        if (solid_mat(material_at(w, v), VG_RAY_SKIP_TRANSPARENT)) fl |= kProbeBuried;
    }
    pwork[j] = make_float4(o[0], o[1], o[2], __uint_as_float(fl));
}

// Path stages of a lane: kNew holds a freshly taken work id whose first ray is not set up yet;
// the walking stages are the kinds of ray the lane traces.
enum : int { kIdle = 0, kNew = 1, kGather = 2, kShadowX = 3, kBounce = 4, kShadowY = 5 };

// Learned-mode arguments of the path kernel.
struct LearnedArgs {
    vg_nrc_query *queries; // two halves of `half` entries (this update's / the previous hits)
    uint32_t half;
    uint32_t train_n;      // training targets traced by this launch (work ids [0, train_n))
    uint32_t tseed;
    vg_nrc_query *tq;      // training batch: queries (aux = bounce-table index)
    float4 *tt;            // ... and targets (rgb, weight)
};

// Longest-first order: the previous update's heavy samples (their ids, and one bit per
// sample) are traced first and skipped in the natural order; this update's are recorded.
// Any order gives the same results; this one starts the long paths early, so the launch
// does not end on a tail of them.
constexpr uint32_t kHeavyIters = 24;  // gather + sun shadow paths (about the top 1 %)
constexpr uint32_t kHeavyIters2 = 48; // 2-bounce paths
struct HeavyArgs {
    uint32_t *ids;   // two halves of `cap` ids; half (updates so far) & 1 is this update's
    uint32_t *mask;  // two halves of `words` words
    uint32_t cap;    // ids each list holds (a longer list is cut: the rest stay in order)
    uint32_t words;
    uint32_t thresh; // walk iterations that make a path heavy
};

// Path state that must outlive a walk, one slot per thread in LDS (structure of arrays, so a
// wave's accesses are bank-conflict free). Only the walk itself stays in registers while the
// lanes march, which keeps the kernel at four waves per SIMD.
struct PathSlots {
    float go[3][kBlock]; // gather ray origin (the probe's) and direction of the lane's sample
    float gd[3][kBlock];
    float so[3][kBlock]; // secondary origin: off the gather hit (training: off the query point)
    float tX[kBlock];    // gather hit distance
    float ndlX[kBlock];  // max(0, n . sun) at the gather hit
    float visX[kBlock];  // its sun visibility (2-bounce paths keep it across the bounce)
    float ndlY[kBlock];  // max(0, n . sun) at the bounce hit
    uint32_t pc[kBlock]; // walk iterations of the lane's sample path so far (~0u: training)
    uint32_t fm[kBlock]; // gather face | gather material << 8 | bounce material << 16
    uint32_t tq[kBlock]; // training lane: index of its query in the previous hit list
    uint32_t tb[kBlock]; // training lane: bounce-table index
};

// Learned mode: finished gather hits wait in a per-wave LDS ring of cache queries and go to
// the global list 64 at a time (one atomic per run of 64, not per refill).
constexpr uint32_t kQStage = 128;
// Work ids of the path kernel, wave-uniform: ids [0, a) (training samples, then the previous
// update's heavy samples) from the global counter ctr[0], then the natural samples [0, n)
// as ids a + sample from the stripes. A wave holds a reservation [next, end) of at least
// `chunk` ids (0: what the refill lacks) and moves to the next stripe when its own is used up.
struct PathQueue {
    uint32_t next, end;
    uint32_t sc;   // current stripe
    uint32_t left; // stripes not yet found used up
    bool a_done, drained;
};

__device__ __forceinline__ bool pq_empty(const PathQueue &pq) {
    return pq.drained && pq.next >= pq.end;
}

// Ids for `need` (<= 64) idle lanes: the idle lane of rank `rank` gets the returned id when
// `want` is set (~0u: none this time). Every lane of the wave calls it together.
__device__ __forceinline__ uint32_t pq_take(PathQueue &pq, uint32_t *ctr, uint32_t a, uint32_t n,
                                            uint32_t slen, uint32_t need, uint32_t rank,
                                            bool want, uint32_t lane, uint32_t chunk) {
    const uint32_t avail = pq.end - pq.next;
    uint32_t b = ~0u, e = 0u;
    if (avail < need && !pq.drained) {
        const uint32_t g = need - avail > chunk ? need - avail : chunk;
        if (!pq.a_done) {
            uint32_t r = 0;
            if (lane == 0) r = atomicAdd(ctr, g);
            r = __builtin_amdgcn_readfirstlane(__shfl(r, 0, 64));
            if (r < a) {
                b = r;
                e = a - r < g ? a : r + g;
            } else {
                pq.a_done = true;
            }
        }
        while (b == ~0u && pq.a_done && pq.left > 0u) {
            uint32_t s0, s1;
            stripe_range(pq.sc, slen, n, s0, s1);
            uint32_t r = 0;
            if (lane == 0) r = atomicAdd(ctr + kGiStripe0 + pq.sc * kGiStripeWords, g);
            r = __builtin_amdgcn_readfirstlane(__shfl(r, 0, 64));
            if (r < s1 - s0) {
                b = a + s0 + r;
                e = a + (s1 - s0 - r < g ? s1 : s0 + r + g);
            } else {
                pq.sc = pq.sc + 1u < (uint32_t)kGiStripes ? pq.sc + 1u : 0u;
                pq.left--;
            }
        }
        if (b == ~0u && pq.a_done && pq.left == 0u) pq.drained = true;
    }
    uint32_t id = ~0u;
    if (want) {
        if (rank < avail) id = pq.next + rank;
        else if (b != ~0u && b + (rank - avail) < e) id = b + (rank - avail);
    }
    if (b != ~0u) {
        const uint32_t nx = b + (need - avail);
        pq.next = nx < e ? nx : e;
        pq.end = e;
    } else {
        pq.next += avail < need ? avail : need;
    }
    return id;
}

template <uint32_t CAP> struct QueryStage {
    float4 q[kBlock / 64][CAP][2];
};

template <bool COUNT, int BOUNCES, bool FOLD, bool MATSPEC>
__global__ __launch_bounds__(kBlock) void k_probe_paths(
    DevWorld w, vg_gi_params p, uint32_t n_samples, uint32_t epoch, uint32_t phase,
    const float4 *__restrict__ pwork, const float4 *__restrict__ dirs,
    const float4 *__restrict__ btab, const float4 *__restrict__ pal, float4 *__restrict__ rad,
    uint8_t *__restrict__ face_out, LearnedArgs la, HeavyArgs ha, uint32_t *ctr, int threshold,
    int tbatch, uint32_t chunk, unsigned long long *counts) {
    world_ready(w);
    // BOUNCES: 1 ambient tail, 2 traced cosine bounce, 3 learned (the cache adds the tail;
    // the lanes' kBounce / kShadowY stages then trace training targets), 4 learned without
    // training targets (an update that takes no training step: no bounce stages compiled in)
    constexpr bool LEARNED = BOUNCES >= 3;
    constexpr bool TRAIN = BOUNCES == 3;
    constexpr bool BOUNCE_STAGES = BOUNCES == 2 || TRAIN;
    __shared__ PathSlots sl;
    __shared__ QueryStage<LEARNED ? kQStage : 1u> qs;
    __shared__ float4 spal[256]; // the albedo palette
    __shared__ float ssinv[3];   // safe_inv of the sun direction, shared by every shadow ray
    for (uint32_t i = threadIdx.x; i < 256u; i += kBlock) spal[i] = pal[i];
    if (threadIdx.x < 3u) ssinv[threadIdx.x] = safe_inv(p.sun[threadIdx.x]);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t me = threadIdx.x;
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t par = LEARNED ? ctr[3] : 0u;
    vg_nrc_query *qcur = la.queries + (size_t)par * la.half;
    const vg_nrc_query *qprev = la.queries + (size_t)(par ^ 1u) * la.half;
    const uint32_t nprev = TRAIN ? ctr[2] : 0u;
    const uint32_t train_n = TRAIN ? la.train_n : 0u;
    const uint32_t n_heavy = ctr[kCtrHeavy + 1] < ha.cap ? ctr[kCtrHeavy + 1] : ha.cap;
    const uint32_t hp = ctr[kCtrUpd] & 1u; // this update's half of the heavy lists
    uint32_t *cur_ids = ha.ids + (size_t)hp * ha.cap, *cur_mask = ha.mask + (size_t)hp * ha.words;
    const uint32_t *prev_ids = ha.ids + (size_t)(hp ^ 1u) * ha.cap;
    const uint32_t *prev_mask = ha.mask + (size_t)(hp ^ 1u) * ha.words;
    uint32_t cnt[VG_IT_COUNT];
    uint32_t nray[3] = {0, 0, 0};
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    const int max_iter = w.dims[0] + w.dims[1] + w.dims[2] + 8;
    const uint32_t N = (uint32_t)p.rays_per_probe;
    int stage = kIdle;
    uint32_t sid = 0; // sample id (training lane: batch index; kNew: the raw work id)
    int it = 0;
    Walk r;
    WalkPhase q;
    HitRec h;
    hit_miss(h);
    phase_reset(q);
    const uint32_t n_upd_probes = n_samples / N;
    const uint32_t slen = ((n_upd_probes + kGiStripes - 1) / kGiStripes) * N; // whole probes
    PathQueue pq;
    pq.sc = blockIdx.x % (uint32_t)kGiStripes; // the stripe of this workgroup's XCD
    pq.left = (uint32_t)kGiStripes;
    pq.a_done = pq.drained = false;
    // A wave whose static reservation holds front samples (the longest paths) takes no other
    // work when the whole front region is reserved statically and other waves are left for
    // the natural samples: its loop trips then carry only its own paths' walks and transitions.
    bool front_only = false;
    {
        // the wave's static first reservation (stripe_waves_below)
        const uint32_t front = train_n + n_heavy;
        const uint32_t nw = gridDim.x * (kBlock / 64);
        const uint32_t lim = front < nw * 64u ? front : nw * 64u;
        const uint32_t nfw = (lim + 63u) / 64u;
        const uint32_t gw = blockIdx.x * (kBlock / 64) + wv;
        if (gw < nfw) {
            pq.next = gw * 64u;
            pq.end = gw * 64u + 64u < lim ? gw * 64u + 64u : lim;
            front_only = lim == front && nfw < nw;
        } else {
            // the front region is all reserved statically: natural samples of the stripe
            uint32_t s0, s1;
            stripe_range(pq.sc, slen, n_samples, s0, s1);
            const uint32_t rank =
                stripe_waves_below(gw, pq.sc) - stripe_waves_below(nfw, pq.sc);
            const uint32_t b = s0 + rank * 64u;
            pq.next = front + (b < s1 ? b : s1);
            pq.end = front + (b + 64u < s1 ? b + 64u : s1);
            pq.a_done = true;
        }
    }
    bool emit = false;  // learned: this lane's finished gather hit still has to be listed
    bool heavy = false; // this lane's finished sample path still has to be recorded heavy
    uint32_t qhead = 0, qn = 0; // the wave's staged queries: ring start and count

    // Finish a sample with a gather hit: radiance toward the probe, given the light lb
    // arriving at the hit from its bounce, and the visible distance.
    auto finish_hit = [&](const float lb[3]) {
        const uint32_t fm = sl.fm[me];
        const float4 a = spal[(fm >> 8) & 0xFFu];
        const float sx = sl.ndlX[me] * sl.visX[me];
        rad[sid] = make_float4(a.x * (p.sun_color[0] * sx + lb[0]),
                               a.y * (p.sun_color[1] * sx + lb[1]),
                               a.z * (p.sun_color[2] * sx + lb[2]), sl.tX[me]);
        face_out[sid] = (uint8_t)(fm & 0xFFu);
    };
    // light leaving the bounce hit toward the gather hit (sun through its shadow ray + ambient)
    auto bounce_light = [&](float visY, float out[3]) {
        const float4 a = spal[(sl.fm[me] >> 16) & 0xFFu];
        const float sy = sl.ndlY[me] * visY;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            out[k] = (&a.x)[k] * (p.sun_color[k] * sy + p.ambient[k] * p.ambient_tail);
    };
    // a cosine bounce off face f from table entry i: direction and its inverses
    // (iv null: the direction only)
    auto bounce_from = [&](uint32_t i, uint32_t f, float d[3], float *iv) {
        const float4 e = btab[i];
        bounce_dir(e, f, d);
        if (iv) bounce_inv(e, btab[VG_GI_BOUNCE_TABLE + i], f, iv);
    };
    auto bounce_ray = [&](float d[3], float *iv) { // plain 2-bounce: the bounce of sample sid
        const uint32_t probe = (sid / N) * (uint32_t)p.update_div + phase;
        bounce_from(bounce_index(p.seed, p.round, probe, sid % N), sl.fm[me] & 0xFFu, d, iv);
    };
    auto train_dir = [&](float d[3], float *iv) {
        bounce_from(sl.tb[me], sl.fm[me] & 0xFFu, d, iv);
    };
    // a training target: the previous list's query, its bounce-table index, the traced light
    auto train_write = [&](const float t[3]) {
        const float4 *src = reinterpret_cast<const float4 *>(qprev + sl.tq[me]);
        const float4 a = src[0], b = src[1];
        float4 *dst = reinterpret_cast<float4 *>(la.tq + sid);
        dst[0] = make_float4(a.x, a.y, a.z, __uint_as_float(sl.fm[me] & 0xFFu));
        dst[1] = make_float4(b.x, b.y, b.z, __uint_as_float(sl.tb[me]));
        la.tt[sid] = make_float4(t[0], t[1], t[2], 1.0f);
    };
    // Learned: a finished hit stays pending in its (idle) lane until the wave's next refill
    // (or its exit), which stages the wave's pending hits in LDS; full runs of 64 (at exit:
    // the rest) are appended to the list with one atomic on ctr[1]. The query (hit point,
    // gather direction, face | material << 8, sample id) is rebuilt from the lane's slot.
    auto stage_queries = [&](bool final) {
        const uint64_t em = __ballot(emit);
        if (emit) {
            const float o[3] = {sl.go[0][me], sl.go[1][me], sl.go[2][me]};
            const float d[3] = {sl.gd[0][me], sl.gd[1][me], sl.gd[2][me]};
            const float t = sl.tX[me];
            const uint32_t k = (qhead + qn + lanes_below(em)) & (kQStage - 1u);
            qs.q[wv][k][0] = make_float4(o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t,
                                         __uint_as_float(sl.fm[me] & 0xFFFFu));
            qs.q[wv][k][1] = make_float4(d[0], d[1], d[2], __uint_as_float(sid));
            emit = false;
        }
        qn += (uint32_t)__popcll(em);
        while (qn >= 64u || (final && qn > 0u)) {
            const uint32_t m = qn < 64u ? qn : 64u;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(ctr + 1, m);
            base = __shfl(base, 0, 64);
            if (lane < m) {
                const uint32_t k = (qhead + lane) & (kQStage - 1u);
                float4 *dst = reinterpret_cast<float4 *>(qcur + base + lane);
                dst[0] = qs.q[wv][k][0];
                dst[1] = qs.q[wv][k][1];
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            qhead = (qhead + m) & (kQStage - 1u);
            qn -= m;
        }
    };

    // Record the wave's finished heavy sample paths (at a refill or the exit, before the
    // lanes take new ids): a slot in this update's list, then the sample's bit.
    auto flush_heavy = [&]() {
        const uint64_t hm = __ballot(heavy);
        if (hm) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(ctr + kCtrHeavy, (uint32_t)__popcll(hm));
            base = __shfl(base, 0, 64);
            if (heavy) {
                const uint32_t k = base + lanes_below(hm);
                if (k < ha.cap) {
                    cur_ids[k] = sid;
                    atomicOr(cur_mask + (sid >> 5), 1u << (sid & 31u));
                }
                heavy = false;
            }
        }
    };

    // The lane's ray of stage `ev` has ended (h holds its result; kNew: a new work id):
    // settle what it decides and set up the next ray of the path, if any. A ray that cannot
    // start (it misses the world window) ends at once without a hit, exactly as a walk that
    // finds nothing, so every ray setup goes through the one walk_init below.
    auto transition = [&]() {
        int ev = stage;
        if (ev >= kGather) {
            const uint32_t c = sl.pc[me];
            if (c != ~0u) sl.pc[me] = c + (uint32_t)it;
        }
        for (;;) {
            float o[3], d[3], iv[3]; // the next ray and the inverses of its direction
            int ns = kIdle;
            bool after_x = false;
            if (ev == kNew) {
                if (TRAIN && sid < train_n) {
                    sl.pc[me] = ~0u;
                    if (nprev == 0u) { // no hits listed yet: a zero-weight target
                        float4 *dst = reinterpret_cast<float4 *>(la.tq + sid);
                        dst[0] = make_float4(0.0f, 0.0f, 0.0f,
                                             __uint_as_float((uint32_t)VG_FACE_NONE));
                        dst[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        la.tt[sid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    } else {
                        uint32_t tqi, tbidx;
                        train_pick(la.tseed, p.round, sid, train_n, nprev, tqi, tbidx);
                        const float4 a = reinterpret_cast<const float4 *>(qprev + tqi)[0];
                        const uint32_t f = __float_as_uint(a.w) & 0xFFu;
                        sl.tq[me] = tqi;
                        sl.tb[me] = tbidx;
                        sl.fm[me] = f;
                        o[0] = a.x + face_n(f, 0) * 0.03f;
                        o[1] = a.y + face_n(f, 1) * 0.03f;
                        o[2] = a.z + face_n(f, 2) * 0.03f;
#pragma unroll
                        for (int k = 0; k < 3; ++k) sl.so[k][me] = o[k];
                        bounce_from(tbidx, f, d, iv);
                        ns = kBounce;
                    }
                } else {
                    // the previous update's heavy samples first, then the others in order
                    const uint32_t k = sid - train_n;
                    bool take;
                    if (k < n_heavy) {
                        sid = prev_ids[k];
                        take = sid < n_samples;
                    } else {
                        sid = k - n_heavy;
                        take = !((prev_mask[sid >> 5] >> (sid & 31u)) & 1u);
                    }
                    sl.pc[me] = 0u;
                    // the probe's origin and flags, the direction and its inverses load
                    // together (and, in the natural order, with the heavy-mask word)
                    const uint32_t sc = sid < n_samples ? sid : 0u;
                    const float4 po = pwork[sc / N];
                    const uint32_t di = (sc % N) * (uint32_t)p.epochs + epoch;
                    const float4 dd = dirs[di];
                    const float4 dv = dirs[N * (uint32_t)p.epochs + di];
                    if (take && !(__float_as_uint(po.w) & kProbeBuried)) {
                        o[0] = po.x;
                        o[1] = po.y;
                        o[2] = po.z;
                        d[0] = dd.x;
                        d[1] = dd.y;
                        d[2] = dd.z;
                        iv[0] = dv.x;
                        iv[1] = dv.y;
                        iv[2] = dv.z;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            sl.go[k][me] = o[k];
                            sl.gd[k][me] = d[k];
                        }
                        ns = kGather;
                    }
                }
            } else if (ev == kGather) {
                if (!h.hit) {
                    // the sky seen along the gather ray: k_probe_reduce evaluates it
                    face_out[sid] = (uint8_t)VG_FACE_NONE;
                } else {
                    const float go[3] = {sl.go[0][me], sl.go[1][me], sl.go[2][me]};
                    const float gd[3] = {sl.gd[0][me], sl.gd[1][me], sl.gd[2][me]};
                    const uint32_t f = h.face;
                    const float ndl = face_ndl(f, p.sun);
                    sl.tX[me] = h.t;
                    sl.fm[me] = f | (h.mat << 8);
                    sl.ndlX[me] = ndl;
                    gi_offset_origin(go, gd, h.t, f, o);
#pragma unroll
                    for (int k = 0; k < 3; ++k) sl.so[k][me] = o[k];
                    if (ndl > 0.0f) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            d[k] = p.sun[k];
                            iv[k] = ssinv[k];
                        }
                        ns = kShadowX;
                    } else {
                        sl.visX[me] = 0.0f;
                        after_x = true;
                    }
                }
            } else if (ev == kShadowX) {
                sl.visX[me] = h.hit ? 0.0f : 1.0f;
                after_x = true;
            } else if (BOUNCE_STAGES && ev == kBounce) {
                // the bounce of a 2-bounce sample, or a learned lane's training path
                if (TRAIN) train_dir(d, nullptr);
                else bounce_ray(d, nullptr);
                if (!h.hit) {
                    float sky[3], t[3];
                    sky_color(p, d, sky);
#pragma unroll
                    for (int k = 0; k < 3; ++k) t[k] = sky[k] * p.gi_sky;
                    if (TRAIN) train_write(t);
                    else finish_hit(t);
                } else {
                    const uint32_t f = h.face;
                    const float ndl = face_ndl(f, p.sun);
                    sl.fm[me] = (sl.fm[me] & 0xFFFFu) | (h.mat << 16);
                    sl.ndlY[me] = ndl;
                    if (ndl > 0.0f) {
                        const float bo[3] = {sl.so[0][me], sl.so[1][me], sl.so[2][me]};
                        gi_offset_origin(bo, d, h.t, f, o);
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            d[k] = p.sun[k];
                            iv[k] = ssinv[k];
                        }
                        ns = kShadowY;
                    } else {
                        float t[3];
                        bounce_light(0.0f, t);
                        if (TRAIN) train_write(t);
                        else finish_hit(t);
                    }
                }
            } else if (BOUNCE_STAGES) { // kShadowY
                float t[3];
                bounce_light(h.hit ? 0.0f : 1.0f, t);
                if (TRAIN) train_write(t);
                else finish_hit(t);
            }
            if (after_x) {
                if (LEARNED) { // the cache supplies the bounce: list the hit
                    const float lb[3] = {0.0f, 0.0f, 0.0f};
                    finish_hit(lb);
                    emit = true;
                } else if (BOUNCES == 2) {
                    bounce_ray(d, iv);
#pragma unroll
                    for (int k = 0; k < 3; ++k) o[k] = sl.so[k][me];
                    ns = kBounce;
                } else {
                    float lb[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) lb[k] = p.ambient[k] * p.ambient_tail;
                    finish_hit(lb);
                }
            }
            if (ns == kIdle) {
                const uint32_t c = sl.pc[me];
                heavy = c != ~0u && c >= ha.thresh;
                stage = kIdle;
                return;
            }
            hit_miss(h);
            if (walk_init(r, w, o, d, iv)) {
                phase_reset(q);
                it = 0;
                stage = ns;
                if (COUNT) nray[ns == kGather ? 0 : (ns == kBounce ? 2 : 1)]++;
                return;
            }
            ev = ns; // could not start: ended without a hit
        }
    };

    // Work ids come from the path queue, `chunk` per atomic (samples of a few probes; 0: what
    // the refill lacks), so a wave refills its idle lanes from its own reservation: training
    // samples and the previous update's heavy samples first, from one counter, then the
    // natural order from the stripes.
    for (;;) {
        uint64_t act = __ballot(stage != kIdle);
        int nact = __popcll(act);
        if (!pq_empty(pq) && nact < threshold && !(front_only && pq.next >= pq.end)) {
            if (LEARNED) stage_queries(false);
            flush_heavy();
            const uint32_t id = pq_take(pq, ctr, train_n + n_heavy, n_samples, slen,
                                        (uint32_t)(64 - nact), lanes_below(~act),
                                        stage == kIdle, lane, chunk);
            if (id != ~0u) {
                sid = id;
                stage = kNew;
            }
            nact = __popcll(__ballot(stage != kIdle));
        }
        if (nact == 0) {
            if (pq_empty(pq) || (front_only && pq.next >= pq.end)) {
                if (LEARNED) stage_queries(true);
                flush_heavy();
                break;
            }
            continue;
        }
        bool pend = false; // this lane's ray has ended; its transition waits for company
        do {
            bool ended = stage == kNew;
            if (stage >= kGather && !pend) {
                const bool shadow = stage == kShadowX || stage == kShadowY;
                bool done = walk_iter<COUNT, FOLD, MATSPEC>(r, q, w, shadow ? p.shadow_tmax : p.gi_dist,
                                                   shadow ? VG_RAY_ANY_HIT
                                                          : VG_RAY_SKIP_TRANSPARENT,
                                                   false, h, cnt);
                ended = done || ++it >= max_iter;
            }
            pend = pend || ended;
            const int npend = __popcll(__ballot(pend));
            if (npend >= tbatch || nact - npend < threshold) {
                if (pend) transition();
                pend = false;
            }
            nact = __popcll(__ballot(stage != kIdle));
        } while (nact >= threshold);
    }
    if (COUNT) {
        flush_counts(cnt, counts);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            unsigned long long v = nray[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0 && v) atomicAdd(&counts[VG_IT_COUNT + k], v);
        }
    }
}

// A group of L lanes per updated probe (L = the sample count rounded up to a power of two,
// 8..64; 64 / L probes per wave): project its samples onto SH-L1 (each lane a fixed strided
// subset, then a butterfly over the group: a fixed order, so deterministic), find the
// nearest hit for relocation, then fold and relocate the 48-float record with one field
// (or a few) per lane.
__global__ __launch_bounds__(kBlock) void k_probe_reduce(DevWorld w, vg_gi_params p, uint32_t n_upd,
                                                        uint32_t epoch, uint32_t phase, uint32_t L,
                                                        const float4 *__restrict__ pwork,
                                                        const float4 *__restrict__ dirs,
                                                        const float4 *__restrict__ rad,
                                                        const uint8_t *__restrict__ face,
                                                        vg_probe *probes, uint32_t *ctr) {
    __shared__ float sums[kBlock / 8][20]; // per group: the 20 projected sums
    // this epoch's gather directions and the sky radiance each of them sees (one evaluation
    // per direction and block instead of one per sample that misses)
    __shared__ float4 sdir[kBlock], ssky[kBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[kCtrUpd] += 1u; // the update's last launch
    const uint32_t N = (uint32_t)p.rays_per_probe;
    const bool staged = N <= kBlock;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t gl = lane & (L - 1u);                 // lane within the group
    const uint32_t grp = threadIdx.x / L;                // group within the block
    const uint32_t j = blockIdx.x * (kBlock / L) + grp;  // updated probe
    const bool live = j < n_upd;
    const uint32_t jj = live ? j : 0u;
    const int64_t idx = (int64_t)jj * p.update_div + phase;
    float *rec = reinterpret_cast<float *>(probes + idx);
    // Every global read of the kernel is issued here, before the direction staging, so they
    // share one round trip: the probe's flags, the record fields this lane folds (f = gl + i
    // L), their staging sums (f + 24), the staging count, and the lane's first sample (its
    // face, and its radiance, which a miss ignores).
    constexpr int kFields = 48 / 8; // fields per lane at the smallest group (L = 8)
    float rv[kFields], rs[kFields];
    float rc = 0.0f;
    uint32_t fs0 = VG_FACE_NONE;
    float4 rad0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t fl = __float_as_uint(pwork[jj].w);
    if (live) {
#pragma unroll
        for (int i = 0; i < kFields; ++i) {
            const uint32_t f = gl + (uint32_t)i * L;
            rv[i] = f < 48u ? rec[f] : 0.0f;
            rs[i] = f < 20u ? rec[24u + f] : 0.0f;
        }
        rc = rec[44];
        if (gl < N) {
            fs0 = face[jj * N + gl];
            rad0 = rad[jj * N + gl];
        }
    }
    if (staged && threadIdx.x < N) {
        const float4 dd = dirs[threadIdx.x * (uint32_t)p.epochs + epoch];
        const float d3[3] = {dd.x, dd.y, dd.z};
        float sky[3];
        sky_color(p, d3, sky);
        sdir[threadIdx.x] = dd;
        ssky[threadIdx.x] = make_float4(sky[0] * p.gi_sky, sky[1] * p.gi_sky, sky[2] * p.gi_sky,
                                        p.gi_dist);
    }
    __syncthreads();
    const bool buried = (fl & kProbeBuried) != 0;
    const float wpp = 12.566371f / (float)N; // 4 pi / N (uniform-sphere MC weight)
    float c[20];
#pragma unroll
    for (int a = 0; a < 20; ++a) c[a] = 0.0f;
    float near_t = 1.0e30f;
    uint32_t near_k = 0xFFFFFFFFu, near_f = VG_FACE_NONE;
    if (live) {
        for (uint32_t k = gl; k < N; k += L) {
            const uint32_t s = jj * N + k;
            const bool first = k == gl;
            const uint32_t fs = first ? fs0 : face[s];
            const float4 rs4 = first ? rad0 : rad[s];
            const float4 dd = staged ? sdir[k] : dirs[k * (uint32_t)p.epochs + epoch];
            if (buried) continue;
            float4 v;
            if (fs == VG_FACE_NONE) { // no gather hit: the sky along the ray, at gi_dist
                if (staged) {
                    v = ssky[k];
                } else {
                    const float d3[3] = {dd.x, dd.y, dd.z};
                    float sky[3];
                    sky_color(p, d3, sky);
                    v = make_float4(sky[0] * p.gi_sky, sky[1] * p.gi_sky, sky[2] * p.gi_sky,
                                    p.gi_dist);
                }
            } else {
                v = rs4;
            }
            const float bs[4] = {0.282095f, 0.488603f * dd.y, 0.488603f * dd.z, 0.488603f * dd.x};
            const float ws[5] = {wpp * v.x, wpp * v.y, wpp * v.z, wpp * v.w, wpp * v.w * v.w};
#pragma unroll
            for (int a = 0; a < 5; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) c[a * 4 + b] += ws[a] * bs[b];
            if (fs != VG_FACE_NONE && v.w < near_t) {
                near_t = v.w;
                near_k = k;
                near_f = fs;
            }
        }
    }
    // butterfly over the group (groups are whole: every lane of a group takes the same path)
    for (uint32_t off = L >> 1; off > 0; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 20; ++a) c[a] += __shfl_xor(c[a], (int)off, 64);
        // nearest hit, ties to the lowest sample (the reference's ascending loop with <)
        const float ot = __shfl_xor(near_t, (int)off, 64);
        const uint32_t ok = __shfl_xor(near_k, (int)off, 64);
        const uint32_t of = __shfl_xor(near_f, (int)off, 64);
        if (ot < near_t || (ot == near_t && ok < near_k)) {
            near_t = ot;
            near_k = ok;
            near_f = of;
        }
    }
    if (gl == 0) {
#pragma unroll
        for (int a = 0; a < 20; ++a) sums[grp][a] = c[a];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (!live) return;
    int32_t g[3];
    probe_cell(p, w, idx, g);
    if (buried) {
        for (uint32_t f = gl; f < 48; f += L) rec[f] = f >= 20 && f < 23 ? (float)g[f - 20] : 0.0f;
        return;
    }
    const bool is_new = (fl & kProbeNew) != 0;
    const bool fold = !is_new && epoch == (uint32_t)p.epochs - 1u;
    // staging count after this update (the fold divides by it, then clears it)
    const float cnt1 = is_new ? 0.0f : rc + 1.0f;
    float n3[3] = {0.0f, 0.0f, 0.0f};
    if (near_k != 0xFFFFFFFFu) {
#pragma unroll
        for (int k = 0; k < 3; ++k) n3[k] = face_n(near_f, k);
    }
#pragma unroll
    for (int i = 0; i < kFields; ++i) {
        const uint32_t f = gl + (uint32_t)i * L;
        if (f >= 48u) break;
        float out;
        if (f < 20) {
            const float cf = sums[grp][f];
            if (is_new) {
                out = cf;
            } else if (fold) {
                const float cnt = cnt1 > 1.0f ? cnt1 : 1.0f;
                const float h = p.hysteresis;
                out = ((rs[i] + cf) / cnt) * (1.0f - h) + rv[i] * h;
            } else {
                out = rv[i];
            }
        } else if (f < 24) {
            out = f < 23 ? (float)g[f - 20] : 1.0f;
        } else if (f < 44) {
            out = (is_new || fold) ? 0.0f : rv[i] + sums[grp][f - 24];
        } else if (f == 44) {
            out = (is_new || fold) ? 0.0f : cnt1;
        } else {
            // relocation (shaders/gi_probes.wgsl:381-392)
            const int k = (int)f - 45;
            float o = is_new ? 0.0f : rv[i];
            if (near_t < 1.2f) {
                const float m = (1.2f - near_t) < 0.35f ? (1.2f - near_t) : 0.35f;
                const float v = o + n3[k] * m;
                o = v < -3.5f ? -3.5f : (v > 3.5f ? 3.5f : v);
            } else if (near_t > 2.5f) {
                o = o * 0.98f;
            }
            out = o;
        }
        rec[f] = out;
    }
}

// Training targets of the learned update's steps after the first (the first step's are
// traced by the path kernel): `batch` samples drawn from the previous learned update's
// hits (train_pick); each traces the hit's cosine bounce ray and the sun shadow ray at the
// bounce hit, like the plain 2-bounce path, and writes the query (aux = bounce-table index)
// and its target (bounce tail rgb, weight 1). With no hits every target gets weight 0.
// Persistent and ray-compacting.
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_nrc_targets(
    DevWorld w, vg_gi_params p, uint32_t batch, uint32_t tseed,
    const vg_nrc_query *__restrict__ qbase, uint32_t half, const uint32_t *gctr,
    const float4 *__restrict__ btab, const float4 *__restrict__ pal,
    vg_nrc_query *__restrict__ tq, float4 *__restrict__ tt, uint32_t *ctr, int threshold,
    unsigned long long *counts) {
    world_ready(w);
    const vg_nrc_query *queries = qbase + (size_t)(gctr[3] ^ 1u) * half;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t cnt[VG_IT_COUNT];
    uint32_t nray[3] = {0, 0, 0};
    if (COUNT)
#pragma unroll
        for (int k = 0; k < VG_IT_COUNT; ++k) cnt[k] = 0;
    const int max_iter = w.dims[0] + w.dims[1] + w.dims[2] + 8;
    const uint32_t nq = gctr[2];
    int stage = kIdle;
    uint32_t j = 0, bidx = 0, fq = 0, mY = 0;
    float qp[3] = {0.0f, 0.0f, 0.0f}, qd[3] = {0.0f, 0.0f, 0.0f};
    float so[3] = {0.0f, 0.0f, 0.0f}, bd[3] = {0.0f, 0.0f, 0.0f};
    float ndlY = 0.0f;
    int it = 0;
    Walk r;
    WalkPhase q;
    HitRec h;
    hit_miss(h);
    phase_reset(q);
    bool exhausted = false;
    auto write = [&](const float lb[3]) {
        float4 *dst = reinterpret_cast<float4 *>(tq + j);
        dst[0] = make_float4(qp[0], qp[1], qp[2], __uint_as_float(fq));
        dst[1] = make_float4(qd[0], qd[1], qd[2], __uint_as_float(bidx));
        tt[j] = make_float4(lb[0], lb[1], lb[2], 1.0f);
        stage = kIdle;
    };
    auto sky_tail = [&]() {
        float sky[3], lb[3];
        sky_color(p, bd, sky);
#pragma unroll
        for (int k = 0; k < 3; ++k) lb[k] = sky[k] * p.gi_sky;
        write(lb);
    };
    auto hit_tail = [&](float visY) {
        const float4 a = pal[mY];
        const float sy = ndlY * visY;
        float lb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            lb[k] = (&a.x)[k] * (p.sun_color[k] * sy + p.ambient[k] * p.ambient_tail);
        write(lb);
    };
    auto start = [&](const float o[3], const float d[3], int st) -> bool {
        hit_miss(h);
        if (!walk_init(r, w, o, d)) return false;
        phase_reset(q);
        it = 0;
        stage = st;
        if (COUNT) nray[st == kBounce ? 2 : 1]++;
        return true;
    };
    for (;;) {
        uint64_t act = __ballot(stage != kIdle);
        int nact = __popcll(act);
        if (!exhausted && nact < threshold) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(ctr, (uint32_t)(64 - nact));
            base = __shfl(base, 0, 64);
            if (base >= batch) {
                exhausted = true;
            } else if (stage == kIdle) {
                const uint32_t id = base + lanes_below(~act);
                if (id < batch) {
                    j = id;
                    if (nq == 0) {
                        float4 *dst = reinterpret_cast<float4 *>(tq + j);
                        dst[0] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float((uint32_t)VG_FACE_NONE));
                        dst[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        tt[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    } else {
                        uint32_t qi;
                        train_pick(tseed, p.round, j, batch, nq, qi, bidx);
                        const float4 *src = reinterpret_cast<const float4 *>(queries + qi);
                        const float4 a = src[0], b = src[1];
                        qp[0] = a.x;
                        qp[1] = a.y;
                        qp[2] = a.z;
                        fq = __float_as_uint(a.w) & 0xFFu;
                        qd[0] = b.x;
                        qd[1] = b.y;
                        qd[2] = b.z;
#pragma unroll
                        for (int k = 0; k < 3; ++k) so[k] = qp[k] + face_n(fq, k) * 0.03f;
                        bounce_dir(btab[bidx], fq, bd);
                        if (!start(so, bd, kBounce)) sky_tail();
                    }
                }
            }
            nact = __popcll(__ballot(stage != kIdle));
        }
        if (nact == 0) {
            if (exhausted) break;
            continue;
        }
        do {
            if (stage != kIdle) {
                const bool shadow = stage == kShadowY;
                bool done = walk_iter<COUNT>(r, q, w, shadow ? p.shadow_tmax : p.gi_dist,
                                             shadow ? VG_RAY_ANY_HIT : VG_RAY_SKIP_TRANSPARENT,
                                             false, h, cnt);
                done = done || ++it >= max_iter;
                if (done) {
                    if (stage == kBounce) {
                        if (!h.hit) {
                            sky_tail();
                        } else {
                            float yo[3];
                            const uint32_t fY = h.face;
                            mY = h.mat;
                            ndlY = face_ndl(fY, p.sun);
                            gi_offset_origin(so, bd, h.t, fY, yo);
                            if (ndlY > 0.0f) {
                                if (!start(yo, p.sun, kShadowY)) hit_tail(1.0f);
                            } else {
                                hit_tail(0.0f);
                            }
                        }
                    } else { // kShadowY
                        hit_tail(h.hit ? 0.0f : 1.0f);
                    }
                }
            }
        } while (__popcll(__ballot(stage != kIdle)) >= threshold);
    }
    if (COUNT) {
        flush_counts(cnt, counts);
#pragma unroll
        for (int k = 1; k < 3; ++k) {
            unsigned long long v = nray[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0 && v) atomicAdd(&counts[VG_IT_COUNT + k], v);
        }
    }
}

// --------------------------------------------------------------------- host --
// ------------------------------------------------------------- probe sampling --
// The probe cache's consumer (shaders/gi_probes.wgsl:69-179 sh_eval_irradiance, sh_eval_scalar,
// sample_probes; shaders/rt_gi.wgsl:128-139 indirect_light in probe mode). Operations are
// plain IEEE fp32 in a fixed order (dot products left to right, no contraction), so the
// oracle's restatement gives the same bits.

constexpr float kGiMaxT = 90.0f;     // GI_MAX_T (shaders/raymarch.wgsl:3821)
constexpr float kGiFadeT = 54.0f;    // GI_MAX_T * 0.6, the fade's start (shaders/rt_gi.wgsl:133)
constexpr float kGiStrength = 1.0f;  // GI_STRENGTH (shaders/rt_gi.wgsl:13)

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ float smoothstep01(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

// Deringed cosine-lobe irradiance / pi of a probe's SH-L1 radiance toward normal n.
__device__ __forceinline__ void sh_irradiance(const float4 *rec, const float n[3], float out[3]) {
    const float a1 = (float)(0.488603 * 0.666667);
    const float lv[3] = {n[1], n[2], n[0]}; // stored order (c1y, c1z, c1x)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float4 sh = rec[c];
        const float l[3] = {sh.y, sh.z, sh.w};
        const float dc = sh.x * 0.282095f;
        const float lin = dot3(l, lv) * a1;
        const float lmax = sqrtf(dot3(l, l)) * a1;
        const float q = 0.95f * dc / (lmax > 1e-6f ? lmax : 1e-6f);
        const float sc = q < 1.0f ? q : 1.0f;
        const float e = dc + sc * lin;
        out[c] = e > 0.0f ? e : 0.0f;
    }
}

// SH-L1 scalar field toward unit direction d (distance moments).
__device__ __forceinline__ float sh_scalar(const float4 c, const float d[3]) {
    return c.x * 0.282095f + c.y * (0.488603f * d[1]) + c.z * (0.488603f * d[2]) +
           c.w * (0.488603f * d[0]);
}

// sample_probes: the 8 probes around p (biased off the surface along n and back along the
// view direction v), smoothstepped trilinear weights, the backface weight and the
// one-tailed Chebyshev visibility bound; probes not gathered for that cell are skipped.
__device__ void sample_probes(const vg_probe *__restrict__ probes, const int32_t grid[3],
                              int32_t spacing, const float pw[3], const float n[3],
                              const float v[3], float out[3]) {
    float pq[3], pg[3], f[3];
    int32_t base[3];
    const float sp = (float)spacing;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pq[k] = (pw[k] + n[k] * 1.0f) - v[k] * 1.0f;
        pg[k] = pq[k] / sp - 0.5f;
        const float fb = floorf(pg[k]);
        base[k] = (int32_t)fb;
        const float fl = pg[k] - fb;
        f[k] = fl * fl * (3.0f - 2.0f * fl);
    }
    float acc[3] = {0.0f, 0.0f, 0.0f};
    float wsum = 0.0f;
    for (int i = 0; i < 8; ++i) {
        const int32_t off[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
        float w = 1.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) w *= off[k] ? f[k] : 1.0f - f[k];
        if (w <= 0.0f) continue;
        int32_t g[3], slot[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            g[k] = base[k] + off[k];
            int32_t m = g[k] % grid[k];
            slot[k] = m < 0 ? m + grid[k] : m;
        }
        const float4 *rec = reinterpret_cast<const float4 *>(
            probes + slot[0] + (int64_t)slot[1] * grid[0] + (int64_t)slot[2] * grid[0] * grid[1]);
        const float4 info = rec[5];
        if (info.w < 0.5f || info.x != (float)g[0] || info.y != (float)g[1] || info.z != (float)g[2])
            continue;
        const float4 st = rec[11];
        const float rel[3] = {st.y, st.z, st.w};
        float dp[3], dn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dp[k] = (((float)g[k] + 0.5f) * sp + rel[k]) - pq[k];
            dn[k] = dp[k] + n[k] * 0.001f;
        }
        const float inv = 1.0f / sqrtf(dot3(dn, dn));
        const float d[3] = {dn[0] * inv, dn[1] * inv, dn[2] * inv};
        float bf = dot3(d, n);
        bf = bf > 0.0f ? bf : 0.0f;
        w = w * (bf * bf + 0.05f);
        const float dist = sqrtf(dot3(dp, dp));
        const float tp[3] = {-d[0], -d[1], -d[2]};
        const float mu = sh_scalar(rec[3], tp);
        float delta = dist - (mu + 0.35f);
        delta = delta > 0.0f ? delta : 0.0f;
        if (delta > 0.0f) {
            const float mu2 = sh_scalar(rec[4], tp);
            float var = mu2 - mu * mu;
            var = var > 0.25f ? var : 0.25f;
            float c = var / (var + delta * delta);
            c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
            w = w * c;
        }
        float e[3];
        sh_irradiance(rec, n, e);
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] = acc[k] + e[k] * w;
        wsum = wsum + w;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = wsum <= 1e-4f ? 0.0f : acc[k] / wsum;
}

// indirect_light in probe mode for every pixel of a traced frame: the hit point is
// camera origin + dir * t (the pixel's primary direction), the normal the hit face's.
// Misses and hits at t >= GI_MAX_T get zero. out4[pixel] = (r, g, b, 0).
__global__ __launch_bounds__(kBlock) void k_frame_indirect(const vg_probe *__restrict__ probes,
                                                           vg_gi_params p, CamParams cam,
                                                           int32_t width, int32_t height,
                                                           const vg_hit *__restrict__ hits,
                                                           float4 *__restrict__ out4) {
    const int64_t pix = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (pix >= (int64_t)width * height) return;
    HitRec h;
    load_hit(hits + pix, h);
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (h.hit && h.t < kGiMaxT) {
        float d[3], pw[3], n[3], e[3];
        primary_dir(cam, (int32_t)(pix % width), (int32_t)(pix / width), d);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pw[k] = cam.origin[k] + d[k] * h.t;
            n[k] = face_n(h.face, k);
        }
        sample_probes(probes, p.grid, p.spacing, pw, n, d, e);
        const float fade = 1.0f - smoothstep01(kGiFadeT, kGiMaxT, h.t);
        const float s = kGiStrength * fade;
        o = make_float4(e[0] * s, e[1] * s, e[2] * s, 0.0f);
    }
    out4[pix] = o;
}

static int check_params(const vg_probes *pr, const vg_gi_params *p, bool learned) {
    if (!p) return fail(VG_E_INVALID, "null params");
    for (int k = 0; k < 3; ++k)
        if (p->grid[k] != pr->grid[k])
            return fail(VG_E_INVALID, "params grid %dx%dx%d does not match the probes' %dx%dx%d",
                        p->grid[0], p->grid[1], p->grid[2], pr->grid[0], pr->grid[1], pr->grid[2]);
    if (p->rays_per_probe != pr->rays || p->epochs != pr->epochs)
        return fail(VG_E_INVALID, "rays/epochs %d/%d do not match the probes' %d/%d",
                    p->rays_per_probe, p->epochs, pr->rays, pr->epochs);
    if (p->spacing <= 0) return fail(VG_E_INVALID, "spacing %d", p->spacing);
    if (p->update_div <= 0) return fail(VG_E_INVALID, "update_div %d", p->update_div);
    if (p->round < 0) return fail(VG_E_INVALID, "round %d", p->round);
    if (!learned && p->bounces != 1 && p->bounces != 2)
        return fail(VG_E_UNSUPPORTED, "bounces %d (1 or 2)", p->bounces);
    return VG_OK;
}

// One probe update: prep, paths, (learned: cache inference), reduce, (learned + train:
// targets and optimizer steps). nrc == nullptr is the plain update.
static int probes_update(vg_probes *pr, const vg_gi_params *p, vg_nrc *nrc, uint32_t nflags,
                         bool count, hipStream_t stream) {
    if (!pr) return fail(VG_E_INVALID, "null probes");
    int st = check_world(pr->world);
    if (st) return st;
    st = check_params(pr, p, nrc != nullptr);
    if (st) return st;
    vg_world *w = pr->world;
    if (nrc && nrc->ctx->device != w->ctx->device)
        return fail(VG_E_INVALID, "the cache and the probes live on different devices");
    const uint32_t div = (uint32_t)p->update_div;
    const uint32_t phase = (uint32_t)p->round % div;
    const uint32_t epoch = ((uint32_t)p->round / div) % (uint32_t)p->epochs;
    const int64_t n_upd64 = pr->total > (int64_t)phase ? (pr->total - phase + div - 1) / div : 0;
    const uint32_t n_upd = (uint32_t)n_upd64;
    if (n_upd == 0) return VG_OK;
    const uint64_t n_samples = (uint64_t)n_upd * (uint64_t)pr->rays;
    unsigned long long *cd = count ? pr->counts : nullptr;
    // heavy-path lists: this update writes one half and reads the other (the previous
    // update's); the device's update count picks them
    HeavyArgs ha;
    ha.ids = pr->heavy;
    ha.mask = pr->hmask;
    ha.words = pr->hwords;
    ha.cap = pr->hcap;
    ha.thresh = (nrc || p->bounces != 2) ? kHeavyIters : kHeavyIters2;
    hipEvent_t *tev = (pr->tev && pr->tn < kTimingCap && !count) ? pr->tev + 5 * pr->tn : nullptr;
    if (tev) {
        pr->tn++;
        HIP_TRY(hipEventRecord(tev[0], stream));
    }
    const dim3 grid((uint32_t)persistent_grid(w->ctx, w->ctx->probe_wpc)), block(kBlock);
    const bool train = nrc && (nflags & VG_NRC_TRAIN) && nrc->cfg.steps > 0;
    const uint32_t train_n = train ? (uint32_t)nrc->tcap : 0u;
    hipLaunchKernelGGL(k_probe_prep, dim3((n_upd + kBlock - 1) / kBlock), dim3(kBlock), 0, stream,
                       w->dw, *p, n_upd, phase, pr->probes, pr->pwork, pr->ctr, nrc ? 1 : 0,
                       pr->hmask, pr->hwords, train_n, ha.cap,
                       grid.x * (uint32_t)(kBlock / 64) * 64u);
    HIP_TRY(hipGetLastError());
    if (tev) HIP_TRY(hipEventRecord(tev[1], stream));
    const int thr = w->ctx->probe_refill;
    // the path launch writes this update's training batch into the half the previous
    // update's training step (possibly still running) does not read
    const int tb = nrc ? nrc->tpar ^ 1 : 0;
    LearnedArgs la;
    la.queries = pr->queries;
    la.half = (uint32_t)pr->ns;
    la.train_n = train_n;
    la.tseed = nrc ? nrc->cfg.seed ^ 0x9E3779B9u : 0u;
    la.tq = nrc ? nrc->tq_base + (size_t)tb * nrc->tcap : nullptr;
    la.tt = nrc ? nrc->tt_base + (size_t)tb * nrc->tcap : nullptr;
#define LAUNCH_GS(C, B, F, S)                                                                  \
    hipLaunchKernelGGL((k_probe_paths<C, B, F, S>), grid, block, 0, stream, w->dw, *p,           \
                       (uint32_t)n_samples, epoch, phase, pr->pwork, pr->dirs, pr->bounce,      \
                       pr->palette, pr->rad, pr->face, la, ha, pr->ctr, thr,                    \
                       w->ctx->probe_tbatch, (uint32_t)w->ctx->probe_chunk, cd)
    // counting runs (one per update at most, outside timed regions) never prefetch
#define LAUNCH_GF(C, B, F)                                                                     \
    do {                                                                                       \
        if (!C && w->ctx->probe_matspec) LAUNCH_GS(C, B, F, !C);                               \
        else LAUNCH_GS(C, B, F, false);                                                        \
    } while (0)
#define LAUNCH_G(C, B)                                                                         \
    do {                                                                                       \
        if (w->dw.fold) LAUNCH_GF(C, B, true);                                                 \
        else LAUNCH_GF(C, B, false);                                                           \
    } while (0)
    // a learned update that takes no training step runs the instantiation without the
    // training stages
    const int mode = nrc ? (la.train_n ? 3 : 4) : p->bounces;
    if (count) {
        if (mode == 4) LAUNCH_G(true, 4);
        else if (mode == 3) LAUNCH_G(true, 3);
        else if (mode == 2) LAUNCH_G(true, 2);
        else LAUNCH_G(true, 1);
    } else {
        if (mode == 4) LAUNCH_G(false, 4);
        else if (mode == 3) LAUNCH_G(false, 3);
        else if (mode == 2) LAUNCH_G(false, 2);
        else LAUNCH_G(false, 1);
    }
#undef LAUNCH_G
#undef LAUNCH_GF
#undef LAUNCH_GS
    HIP_TRY(hipGetLastError());
    if (tev) HIP_TRY(hipEventRecord(tev[2], stream));
    if (nrc) {
        st = nrc_infer_launch(nrc, pr->queries, pr->ns, pr->ctr + 3, pr->ctr + 1,
                              (uint32_t)n_samples, nullptr, pr->rad, pr->palette, stream);
        if (st) return st;
        if (train) {
            nrc->tpar = tb;
            nrc->tq = la.tq;
            nrc->tt = la.tt;
        }
    }
    // one training step overlaps the next update (it needs nothing the next update's prep and
    // path launch write); more steps, or a counting run, stay on the caller's stream
    const bool overlap = train && nrc->cfg.steps == 1 && !count;
    if (overlap) HIP_TRY(hipEventRecord(nrc->ev_ready, stream));
    if (tev) HIP_TRY(hipEventRecord(tev[3], stream));
    // lanes per probe in the reduction: half the sample count's power of two (two samples per
    // lane before the butterfly, which halves it), 8..64
    uint32_t lpp = 8;
    while (lpp < 64 && 2u * lpp < (uint32_t)pr->rays) lpp <<= 1;
    const uint32_t ppb = kBlock / lpp;
    hipLaunchKernelGGL(k_probe_reduce, dim3((n_upd + ppb - 1) / ppb), dim3(kBlock), 0, stream,
                       w->dw, *p, n_upd, epoch, phase, lpp, pr->pwork, pr->dirs, pr->rad,
                       pr->face, pr->probes, pr->ctr);
    HIP_TRY(hipGetLastError());
    if (tev) HIP_TRY(hipEventRecord(tev[4], stream));
    if (overlap) {
        HIP_TRY(hipStreamWaitEvent(nrc->tstream, nrc->ev_ready, 0));
        st = nrc_train_launch(nrc, nrc->tq, nrc->tt, (uint32_t)nrc->tcap, true, nrc->tstream);
        if (st) return st;
        HIP_TRY(hipEventRecord(nrc->ev_trained, nrc->tstream));
        if (nrc_capturing(stream)) {
            // under graph capture the step is joined back into the captured stream, so the
            // graph holds the whole update (and nothing outside it waits for the step)
            HIP_TRY(hipStreamWaitEvent(stream, nrc->ev_trained, 0));
            nrc->pending = false;
        } else {
            nrc->pending = true;
        }
    } else if (train) {
        for (int s = 0; s < nrc->cfg.steps; ++s) {
            if (s > 0) { // step 0's targets came with the path launch
                const uint32_t tseed = nrc->cfg.seed ^ (0x9E3779B9u * (uint32_t)(s + 1));
                if (count)
                    hipLaunchKernelGGL(k_nrc_targets<true>, grid, block, 0, stream, w->dw, *p,
                                       (uint32_t)nrc->tcap, tseed, pr->queries, (uint32_t)pr->ns,
                                       pr->ctr, pr->bounce, pr->palette, nrc->tq, nrc->tt,
                                       pr->ctr + 4 + (s - 1), thr, cd);
                else
                    hipLaunchKernelGGL(k_nrc_targets<false>, grid, block, 0, stream, w->dw, *p,
                                       (uint32_t)nrc->tcap, tseed, pr->queries, (uint32_t)pr->ns,
                                       pr->ctr, pr->bounce, pr->palette, nrc->tq, nrc->tt,
                                       pr->ctr + 4 + (s - 1), thr, cd);
                HIP_TRY(hipGetLastError());
            }
            st = nrc_train_launch(nrc, nrc->tq, nrc->tt, (uint32_t)nrc->tcap, true, stream);
            if (st) return st;
        }
    }
    return VG_OK;
}

static int probes_update_counts(vg_probes *pr, const vg_gi_params *p, vg_nrc *nrc,
                                uint32_t nflags, uint64_t rays_host[3], uint64_t counts_host[],
                                uint64_t *queries_host) {
    if (!pr) return fail(VG_E_INVALID, "null probes");
    int st = set_device(pr->world);
    if (st) return st;
    HIP_TRY(hipMemset(pr->counts, 0, kGiCounters * sizeof(unsigned long long)));
    st = probes_update(pr, p, nrc, nflags, true, nullptr);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    unsigned long long c[kGiCounters];
    HIP_TRY(hipMemcpy(c, pr->counts, sizeof(c), hipMemcpyDeviceToHost));
    if (counts_host)
        for (int k = 0; k < VG_IT_COUNT; ++k) counts_host[k] = c[k];
    if (rays_host)
        for (int k = 0; k < 3; ++k) rays_host[k] = c[VG_IT_COUNT + k];
    if (queries_host) {
        uint32_t nq = 0;
        HIP_TRY(hipMemcpy(&nq, pr->ctr + 1, sizeof(nq), hipMemcpyDeviceToHost));
        *queries_host = nq;
    }
    return VG_OK;
}

static void free_timing(vg_probes *p) {
    if (!p->tev) return;
    for (int i = 0; i < 5 * kTimingCap; ++i)
        if (p->tev[i]) (void)hipEventDestroy(p->tev[i]);
    delete[] p->tev;
    p->tev = nullptr;
    p->tn = 0;
}

static void free_probes(vg_probes *p) {
    void *ptrs[] = {p->probes, p->dirs,    p->bounce, p->palette, p->pwork, p->rad,
                    p->face,   p->queries, p->ctr,    p->counts,  p->heavy, p->hmask};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    free_timing(p);
}

} // namespace vg

using namespace vg;

extern "C" {

int vg_probes_create(vg_world *world, const vg_gi_params *p, vg_probes **out) {
    if (!world || !p || !out) return fail(VG_E_INVALID, "null argument");
    *out = nullptr;
    int st = set_device(world);
    if (st) return st;
    for (int k = 0; k < 3; ++k)
        if (p->grid[k] <= 0) return fail(VG_E_INVALID, "grid[%d] = %d", k, p->grid[k]);
    if (p->rays_per_probe <= 0 || p->rays_per_probe > 65536 || p->epochs <= 0 ||
        (int64_t)p->rays_per_probe * p->epochs > (1 << 24))
        return fail(VG_E_INVALID, "rays_per_probe %d, epochs %d", p->rays_per_probe, p->epochs);
    const int64_t total = (int64_t)p->grid[0] * p->grid[1] * p->grid[2];
    if (total > (1ll << 26) || total * p->rays_per_probe > 0xF0000000ll)
        return fail(VG_E_INVALID, "probe grid too large");
    vg_probes *pr = new vg_probes();
    std::memset(pr, 0, sizeof(*pr));
    pr->world = world;
    pr->device = world->ctx->device;
    pr->ns = (size_t)total * p->rays_per_probe;
    for (int k = 0; k < 3; ++k) pr->grid[k] = p->grid[k];
    pr->rays = p->rays_per_probe;
    pr->epochs = p->epochs;
    pr->total = total;
    const size_t nd = (size_t)p->rays_per_probe * p->epochs;
    const size_t ns = (size_t)total * p->rays_per_probe;
    pr->hwords = (uint32_t)((ns + 31) / 32);
    pr->hcap = (uint32_t)(ns / 32 > 4096 ? ns / 32 : 4096);
    bool ok = hipMalloc((void **)&pr->probes, (size_t)total * sizeof(vg_probe)) == hipSuccess &&
              hipMalloc((void **)&pr->dirs, 2 * nd * sizeof(float4)) == hipSuccess &&
              hipMalloc((void **)&pr->bounce, 2 * VG_GI_BOUNCE_TABLE * sizeof(float4)) ==
                  hipSuccess &&
              hipMalloc((void **)&pr->palette, 256 * sizeof(float4)) == hipSuccess &&
              hipMalloc((void **)&pr->pwork, (size_t)total * sizeof(float4)) == hipSuccess &&
              hipMalloc((void **)&pr->rad, ns * sizeof(float4)) == hipSuccess &&
              hipMalloc((void **)&pr->face, ns) == hipSuccess &&
              hipMalloc((void **)&pr->queries, 2 * ns * sizeof(vg_nrc_query)) == hipSuccess &&
              hipMalloc((void **)&pr->ctr, kGiCtrWords * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void **)&pr->heavy, 2 * (size_t)pr->hcap * sizeof(uint32_t)) ==
                  hipSuccess &&
              hipMalloc((void **)&pr->hmask, 2 * (size_t)pr->hwords * sizeof(uint32_t)) ==
                  hipSuccess &&
              hipMalloc((void **)&pr->counts, kGiCounters * sizeof(unsigned long long)) ==
                  hipSuccess;
    if (!ok) {
        free_probes(pr);
        delete pr;
        return fail(VG_E_NOMEM, "probe storage for %lld probes x %d rays", (long long)total,
                    p->rays_per_probe);
    }
    // each table is followed by the walk's per-component inverses of its entries (safe_inv,
    // in host IEEE single precision: the device's correctly rounded divide gives the same)
    auto add_inverses = [](std::vector<float> &t, size_t n) {
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k) {
                const float x = t[4 * i + k];
                t[4 * (n + i) + k] = k < 3 ? (std::fabs(x) < 1e-8f ? 1e30f : 1.0f / x) : 0.0f;
            }
    };
    std::vector<float> tab(2 * nd * 4);
    vg_gi_directions((int32_t)nd, tab.data());
    add_inverses(tab, nd);
    std::vector<float> bt(2 * VG_GI_BOUNCE_TABLE * 4), pal(256 * 4);
    vg_gi_bounce_table(bt.data());
    add_inverses(bt, VG_GI_BOUNCE_TABLE);
    vg_gi_default_palette(pal.data());
    hipError_t e = hipMemcpy(pr->dirs, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(pr->bounce, bt.data(), bt.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(pr->palette, pal.data(), pal.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(pr->probes, 0, (size_t)total * sizeof(vg_probe));
    if (e == hipSuccess) e = hipMemset(pr->ctr, 0, kGiCtrWords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(pr->hmask, 0, 2 * (size_t)pr->hwords * sizeof(uint32_t));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_probes(pr);
        delete pr;
        return fail(VG_E_DEVICE, "probe setup: %s", hipGetErrorString(e));
    }
    *out = pr;
    return VG_OK;
}

int vg_probes_destroy(vg_probes *pr) {
    if (!pr) return VG_OK;
    (void)hipSetDevice(pr->device);
    (void)hipDeviceSynchronize();
    free_probes(pr);
    delete pr;
    return VG_OK;
}

int vg_frame_indirect(vg_probes *pr, const vg_gi_params *p, const vg_camera *camera,
                      int32_t width, int32_t height, const vg_hit *hits, float *out4,
                      void *stream) {
    if (!pr || !p || !camera) return fail(VG_E_INVALID, "null argument");
    int st = check_world(pr->world);
    if (st) return st;
    for (int k = 0; k < 3; ++k)
        if (p->grid[k] != pr->grid[k]) return fail(VG_E_INVALID, "params grid differs from the probes'");
    if (p->spacing <= 0) return fail(VG_E_INVALID, "spacing %d", p->spacing);
    if (width <= 0 || height <= 0) return fail(VG_E_INVALID, "size %dx%d", width, height);
    if ((int64_t)width * height > (1ll << 31)) return fail(VG_E_INVALID, "frame too large");
    // the pixel rays are rebuilt from the camera: its resolution is the frame's
    if (camera->resolution[0] != (float)width || camera->resolution[1] != (float)height)
        return fail(VG_E_INVALID, "camera resolution %gx%g does not match %dx%d",
                    camera->resolution[0], camera->resolution[1], width, height);
    if (!hits || !out4) return fail(VG_E_INVALID, "null hits/out");
    if (((uintptr_t)out4 & 15u) != 0) return fail(VG_E_INVALID, "out4 must be 16-byte aligned");
    if (((uintptr_t)hits & 3u) != 0) return fail(VG_E_INVALID, "hits must be 4-byte aligned");
    const int64_t n = (int64_t)width * height;
    hipLaunchKernelGGL(k_frame_indirect, dim3((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, pr->probes, *p, cam_params(camera), width, height, hits,
                       reinterpret_cast<float4 *>(out4));
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

int vg_probes_set_palette(vg_probes *pr, const float *rgba) {
    if (!pr || !rgba) return fail(VG_E_INVALID, "null argument");
    int st = set_device(pr->world);
    if (st) return st;
    HIP_TRY(hipMemcpy(pr->palette, rgba, 256 * sizeof(float4), hipMemcpyHostToDevice));
    return VG_OK;
}

int vg_probes_reset(vg_probes *pr, void *stream) {
    if (!pr) return fail(VG_E_INVALID, "null probes");
    int st = set_device(pr->world);
    if (st) return st;
    HIP_TRY(hipMemsetAsync(pr->probes, 0, (size_t)pr->total * sizeof(vg_probe),
                           (hipStream_t)stream));
    return VG_OK;
}

int vg_probes_data(vg_probes *pr, vg_probe **out) {
    if (!pr || !out) return fail(VG_E_INVALID, "null argument");
    *out = pr->probes;
    return VG_OK;
}

int vg_probes_download(vg_probes *pr, vg_probe *host) {
    if (!pr || !host) return fail(VG_E_INVALID, "null argument");
    int st = set_device(pr->world);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host, pr->probes, (size_t)pr->total * sizeof(vg_probe),
                      hipMemcpyDeviceToHost));
    return VG_OK;
}

int vg_probes_upload(vg_probes *pr, const vg_probe *host) {
    if (!pr || !host) return fail(VG_E_INVALID, "null argument");
    int st = set_device(pr->world);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(pr->probes, host, (size_t)pr->total * sizeof(vg_probe),
                      hipMemcpyHostToDevice));
    return VG_OK;
}

int vg_probes_update(vg_probes *pr, const vg_gi_params *p, void *stream) {
    return probes_update(pr, p, nullptr, 0u, false, (hipStream_t)stream);
}

int vg_probes_update_learned(vg_probes *pr, const vg_gi_params *p, vg_nrc *nrc, uint32_t flags,
                             void *stream) {
    if (!nrc) return fail(VG_E_INVALID, "null cache");
    return probes_update(pr, p, nrc, flags, false, (hipStream_t)stream);
}

int vg_autotune_probes(vg_probes *pr, const vg_gi_params *p, vg_nrc *nrc, uint32_t flags,
                       float *best_ms) {
    if (!pr) return fail(VG_E_INVALID, "null probes");
    int st = check_world(pr->world);
    if (st) return st;
    vg_ctx *c = pr->world->ctx;
    // With training, a learned update's step overlaps the next update, so a candidate is
    // timed on two updates back to back (the steady state) and the span closes on the
    // second one's step; the reported time is per update.
    const bool pair = nrc && (flags & VG_NRC_TRAIN) && nrc->cfg.steps == 1;
    auto run = [&](hipStream_t s) {
        int r = probes_update(pr, p, nrc, flags, false, s);
        if (r == VG_OK && pair) r = probes_update(pr, p, nrc, flags, false, s);
        return r == VG_OK && nrc ? nrc_join(nrc, s) : r;
    };
    // for each material prefetch setting: waves per CU x refill x ids per atomic, then the
    // transition batch at the chosen point; the faster of the two settings is kept
    static const int kTbatch[] = {1, 4, 8, 16};
    float best_all = INFINITY;
    int keep[5] = {c->probe_wpc, c->probe_refill, c->probe_chunk, c->probe_tbatch, 0};
    for (int ms = 0; ms < 2 && st == VG_OK; ++ms) {
        c->probe_matspec = ms;
        float b = INFINITY;
        st = autotune(c, &c->probe_wpc, &c->probe_refill, run, &b, &c->probe_chunk);
        const int wpc = c->probe_wpc, rf = c->probe_refill;
        if (st == VG_OK)
            st = autotune_lists(c, &c->probe_wpc, TuneList{&wpc, 1}, &c->probe_refill,
                                TuneList{&rf, 1}, &c->probe_tbatch, TuneList{kTbatch, 4}, run,
                                &b);
        if (st == VG_OK && b < best_all) {
            best_all = b;
            keep[0] = c->probe_wpc;
            keep[1] = c->probe_refill;
            keep[2] = c->probe_chunk;
            keep[3] = c->probe_tbatch;
            keep[4] = ms;
        }
    }
    c->probe_wpc = keep[0];
    c->probe_refill = keep[1];
    c->probe_chunk = keep[2];
    c->probe_tbatch = keep[3];
    c->probe_matspec = keep[4];
    if (best_ms) *best_ms = pair ? 0.5f * best_all : best_all;
    return st;
}

int vg_probes_update_learned_counts(vg_probes *pr, const vg_gi_params *p, vg_nrc *nrc,
                                    uint32_t flags, uint64_t rays_host[3],
                                    uint64_t *queries_host, uint64_t counts_host[VG_IT_COUNT]) {
    if (!nrc) return fail(VG_E_INVALID, "null cache");
    return probes_update_counts(pr, p, nrc, flags, rays_host, counts_host, queries_host);
}

int vg_probes_set_timing(vg_probes *pr, int on) {
    if (!pr) return fail(VG_E_INVALID, "null probes");
    int st = set_device(pr->world);
    if (st) return st;
    if (!on) {
        HIP_TRY(hipDeviceSynchronize());
        free_timing(pr);
        return VG_OK;
    }
    if (!pr->tev) {
        pr->tev = new hipEvent_t[5 * kTimingCap]();
        for (int i = 0; i < 5 * kTimingCap; ++i)
            if (hipEventCreate(&pr->tev[i]) != hipSuccess) {
                free_timing(pr);
                return fail(VG_E_DEVICE, "timing events");
            }
    }
    pr->tn = 0;
    return VG_OK;
}

int vg_probes_timing(vg_probes *pr, float ms_out[4], int32_t *n_out) {
    if (!pr || !ms_out) return fail(VG_E_INVALID, "null argument");
    if (!pr->tev) return fail(VG_E_INVALID, "timing is off (vg_probes_set_timing)");
    int st = set_device(pr->world);
    if (st) return st;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int n = pr->tn;
    for (int u = 0; u < n; ++u) {
        hipEvent_t *e = pr->tev + 5 * u;
        HIP_TRY(hipEventSynchronize(e[4]));
        for (int k = 0; k < 4; ++k) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, e[k], e[k + 1]));
            acc[k] += ms;
        }
    }
    for (int k = 0; k < 4; ++k) ms_out[k] = n ? (float)(acc[k] / n) : 0.0f;
    if (n_out) *n_out = n;
    pr->tn = 0;
    return VG_OK;
}

int vg_probes_update_counts(vg_probes *pr, const vg_gi_params *p, uint64_t rays_host[3],
                            uint64_t counts_host[VG_IT_COUNT]) {
    return probes_update_counts(pr, p, nullptr, 0u, rays_host, counts_host, nullptr);
}

} // extern "C"
