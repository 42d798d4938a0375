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
    int rays_wpc;        // persistent batched kernel: waves per CU
    int rays_refill;
};

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
    size_t beam_cap;                 // entries of beam_t
};

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

// Host-side normalize_or_zero for the sun direction (same expression as the device).
inline bool host_normalize(const float in[3], float out[3]) {
    float len = sqrtf(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
    float rcp = 1.0f / len;
    if (!(std::isfinite(rcp) && rcp > 0.0f)) return false;
    for (int k = 0; k < 3; ++k) out[k] = in[k] * rcp;
    return !(out[0] == 0.0f && out[1] == 0.0f && out[2] == 0.0f);
}

inline int persistent_grid(const vg_ctx *c, int wpc) {
    int waves = c->cus * (wpc > 0 ? wpc : 1);
    return waves / (kBlock / 64) > 0 ? waves / (kBlock / 64) : 1;
}

} // namespace vg
