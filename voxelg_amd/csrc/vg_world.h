// Host-side state of a context and a world on the device, shared by the engine's
// source files. Not part of the ABI.
#pragma once

#include "vg_walk.h"


struct vg_ctx {
    int device;
    int cus;             // compute units of the device
    int schedule;        // VG_SCHEDULE_*
    int frame_wpc;       // persistent frame kernel: waves per CU
    int frame_refill;    // ... refill when fewer lanes of a wave are busy
    int frame_stripes;   // ... work stripes, each with its own counter (1..64)
    int frame_bws;       // ... log2 of the pixel-block width (blocks of 2^bws x 64 / 2^bws)
    int rays_wpc;        // persistent batched kernel: waves per CU
    int rays_refill;
    int probe_wpc;       // persistent probe path kernel: waves per CU
    int probe_refill;
    int probe_chunk;     // ... work ids per atomic (0: exactly what a refill lacks)
    int probe_tbatch;    // ... lanes whose rays ended that set up their next rays together
    int probe_matspec;   // ... 1: each walk step loads its voxel's material with the pyramid words
};

struct vg_edit_state; // sphere-edit buffers (vg_edit.hip), created on first use

struct vg_world {
    vg_ctx *ctx;
    vg::Dims d;
    vg::DevWorld dw;
    bool uploaded;
    uint64_t *l4, *chunk, *tile, *occ;
    uint8_t *tile_uni, *brick_uni, *mat;
    void *scratch;                   // one ray + one hit for vg_raycast
    unsigned long long *counts_dev;  // VG_IT_COUNT totals for the *_counts entry points
    uint32_t *work;                  // persistent kernels' ray counter
    float *beam_t;                   // beam start t per 8x8 pixel block
    int32_t *ytop_dev;               // DevWorld::ytop_p: the y bound of the occupied voxels
    size_t beam_cap;                 // entries of beam_t
    vg_edit_state *edit;
    uint32_t *wlist;                 // water pixels of the last vg_trace_water_secondary
    uint32_t *wlist_n;               // ... their count (device)
    size_t wlist_cap;
};

void vg_edit_state_free(vg_world *w);

namespace vg {

inline int set_device(const vg_world *w) {
    HIP_TRY(hipSetDevice(w->ctx->device));
    return VG_OK;
}

inline int check_world(const vg_world *w) {
    if (!w) return fail(VG_E_INVALID, "null world");
    if (!w->uploaded) return fail(VG_E_STATE, "world has not been uploaded");
    return set_device(w);
}

// The kernels' view of a vg_camera (CameraUniform order, src/camera.rs:25-166).
inline CamParams cam_params(const vg_camera *c) {
    CamParams p;
    for (int k = 0; k < 3; ++k) {
        p.origin[k] = c->origin[k];
        p.forward[k] = c->forward[k];
        p.right[k] = c->right[k];
        p.up[k] = c->up[k];
    }
    p.tan_half_fov = c->tan_half_fov;
    p.res[0] = c->resolution[0];
    p.res[1] = c->resolution[1];
    p.jitter[0] = c->jitter[0];
    p.jitter[1] = c->jitter[1];
    return p;
}

// Host-side normalize_or_zero for the sun direction (same expression as the device).
inline bool host_normalize(const float in[3], float out[3]) {
    float len = sqrtf(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
    float rcp = 1.0f / len;
    if (!(std::isfinite(rcp) && rcp > 0.0f)) return false;
    for (int k = 0; k < 3; ++k) out[k] = in[k] * rcp;
    return !(out[0] == 0.0f && out[1] == 0.0f && out[2] == 0.0f);
}

// Time `launch` (bound to its arguments) over every (waves per CU, refill threshold[, third
// option]) of the candidate lists and keep the fastest in the options: each candidate runs
// once to warm up, then `reps` times between events on one stream; its score is the fastest
// repetition. With a third option (work ids per atomic of the probe kernels, work stripes of
// the frame kernel) its candidates form a third dimension of the sweep.
struct TuneList {
    const int *v;
    int n;
};

template <class L>
inline int autotune_lists(vg_ctx *c, int *wpc_opt, TuneList wl, int *refill_opt, TuneList rl,
                          int *third_opt, TuneList tl, L launch, float *best_ms) {
    const int n_third = third_opt ? tl.n : 1;
    const int reps = 3;
    hipStream_t stream;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int old_schedule = c->schedule;
    c->schedule = VG_SCHEDULE_PERSISTENT;
    float best = INFINITY;
    int bw = *wpc_opt, br = *refill_opt, bt = third_opt ? *third_opt : 0, st = VG_OK;
    const int total = wl.n * rl.n * n_third;
    for (int ci = 0; ci < total && st == VG_OK; ++ci) {
        const int wpc = wl.v[ci % wl.n], rf = rl.v[(ci / wl.n) % rl.n];
        *wpc_opt = wpc;
        *refill_opt = rf;
        if (third_opt) *third_opt = tl.v[ci / (wl.n * rl.n)];
        st = launch(stream);
        if (st) break;
        float ms_min = INFINITY;
        for (int i = 0; i < reps && st == VG_OK; ++i) {
            float ms = 0.0f;
            if (hipEventRecord(e0, stream) != hipSuccess) st = fail(VG_E_DEVICE, "event");
            if (!st) st = launch(stream);
            if (!st && hipEventRecord(e1, stream) != hipSuccess) st = fail(VG_E_DEVICE, "event");
            if (!st && hipEventSynchronize(e1) != hipSuccess) st = fail(VG_E_DEVICE, "sync");
            if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = fail(VG_E_DEVICE, "elapsed");
            ms_min = ms < ms_min ? ms : ms_min;
        }
        if (st) break;
        if (ms_min < best) {
            best = ms_min;
            bw = wpc;
            br = rf;
            if (third_opt) bt = *third_opt;
        }
    }
    *wpc_opt = bw;
    *refill_opt = br;
    if (third_opt) *third_opt = bt;
    c->schedule = old_schedule;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(stream);
    if (st) return st;
    if (best_ms) *best_ms = best;
    return VG_OK;
}

// The batched-query and probe sweeps: waves per CU {4..32} x refill {4..48}, and with chunk_opt
// the probe kernels' work ids per atomic {0, 16..256}.
template <class L>
inline int autotune(vg_ctx *c, int *wpc_opt, int *refill_opt, L launch, float *best_ms,
                    int *chunk_opt = nullptr) {
    static const int kWpc[] = {4, 8, 12, 16, 20, 24, 32};
    static const int kRefill[] = {4, 8, 12, 16, 24, 32, 48};
    static const int kChunk[] = {0, 16, 32, 64, 128, 256};
    return autotune_lists(c, wpc_opt, TuneList{kWpc, 7}, refill_opt, TuneList{kRefill, 7},
                          chunk_opt, TuneList{kChunk, 6}, launch, best_ms);
}


// Dirty-block list (ids of the set bits of a frame block mask, count on the device).
int edit_block_list(vg_world *w, const uint32_t *mask, uint32_t nblocks, hipStream_t stream,
                    const uint32_t **list, const uint32_t **list_n);

inline int persistent_grid(const vg_ctx *c, int wpc) {
    int waves = c->cus * (wpc > 0 ? wpc : 1);
    return waves / (kBlock / 64) > 0 ? waves / (kBlock / 64) : 1;
}

} // namespace vg
