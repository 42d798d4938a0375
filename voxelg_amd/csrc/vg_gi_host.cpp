// Host-side inputs of the GI probe update: the reference's default constants, the
// lighting derived from a sun direction, the albedo palette, and the two direction
// tables the device kernels read (spherical-Fibonacci gather set, cosine-weighted
// bounce set). All of it is computed once on the CPU and uploaded.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "vg_internal.h"

namespace vg {

// WGSL smoothstep / mix / clamp in f32 (the shaders compute the lighting in f32).
static float smooth(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}
static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// Van der Corput radical inverse in base 2 (the second Hammersley coordinate).
static double radical_inverse2(uint32_t i) {
    i = (i << 16) | (i >> 16);
    i = ((i & 0x55555555u) << 1) | ((i & 0xAAAAAAAAu) >> 1);
    i = ((i & 0x33333333u) << 2) | ((i & 0xCCCCCCCCu) >> 2);
    i = ((i & 0x0F0F0F0Fu) << 4) | ((i & 0xF0F0F0F0u) >> 4);
    i = ((i & 0x00FF00FFu) << 8) | ((i & 0xFF00FF00u) >> 8);
    return (double)i * (1.0 / 4294967296.0);
}

} // namespace vg

using namespace vg;

extern "C" {

int vg_gi_params_default(const float sun[3], vg_gi_params *o) {
    if (!sun || !o) return fail(VG_E_INVALID, "null sun/params");
    float len = std::sqrt(sun[0] * sun[0] + sun[1] * sun[1] + sun[2] * sun[2]);
    if (!(len > 0.0f) || !std::isfinite(len)) return fail(VG_E_INVALID, "zero sun direction");
    std::memset(o, 0, sizeof(*o));
    // src/world_dims.rs:44-48; shaders/gi_probes.wgsl:248-272; shaders/rt_gi.wgsl:12-14
    o->grid[0] = 64;
    o->grid[1] = 32;
    o->grid[2] = 64;
    o->spacing = 8;
    o->rays_per_probe = 8;
    o->epochs = 8;
    o->update_div = 8;
    o->round = 0;
    o->bounces = 1;
    o->seed = 0x9E3779B9u;
    o->flags = VG_GI_SKIP_BURIED;
    o->hysteresis = 0.6f;
    o->gi_dist = 40.0f;
    o->shadow_tmax = 480.0f; // SHADOW_MAX_DIST, shaders/raymarch.wgsl:4517
    o->gi_sky = 0.32f;
    o->ambient_tail = 0.35f;
    const float rcp = 1.0f / len;
    float s[3] = {sun[0] * rcp, sun[1] * rcp, sun[2] * rcp};
    for (int k = 0; k < 3; ++k) o->sun[k] = s[k];
    // sun_intensity / sun_color (shaders/common.wgsl:68-83)
    const float day_t = smooth(-0.05f, 0.10f, s[1]);
    const float h = s[1] < 0.0f ? 0.0f : (s[1] > 1.0f ? 1.0f : s[1]);
    const float ember[3] = {1.55f, 0.52f, 0.16f}, gold[3] = {1.38f, 0.88f, 0.44f},
                mid[3] = {1.10f, 1.02f, 0.92f};
    const float a = smooth(0.02f, 0.16f, h), b = smooth(0.16f, 0.55f, h);
    for (int k = 0; k < 3; ++k) o->sun_color[k] = mixf(mixf(ember[k], gold[k], a), mid[k], b) * day_t;
    // ambient_color (shaders/raymarch.wgsl:2870-2876)
    const float aday[3] = {0.30f, 0.42f, 0.58f}, anight[3] = {0.04f, 0.05f, 0.10f};
    for (int k = 0; k < 3; ++k) o->ambient[k] = mixf(anight[k], aday[k], day_t);
    // sky_color's sun-dependent terms (shaders/raymarch.wgsl:2532-2579)
    const float dz[3] = {0.22f, 0.50f, 0.95f}, dh[3] = {0.78f, 0.86f, 0.95f};
    const float nz[3] = {0.005f, 0.010f, 0.035f}, nh[3] = {0.030f, 0.040f, 0.090f};
    for (int k = 0; k < 3; ++k) {
        o->sky_zenith[k] = mixf(nz[k], dz[k], day_t);
        o->sky_horizon[k] = mixf(nh[k], dh[k], day_t);
    }
    o->day_t = day_t;
    o->dusk_w = smooth(-0.08f, 0.18f, s[1]) * (1.0f - smooth(0.05f, 0.40f, s[1]));
    o->halo = day_t * 0.45f;
    const float lxz = std::sqrt(s[0] * s[0] + s[2] * s[2]);
    if (lxz > 0.0f) {
        o->sun_xz[0] = s[0] / lxz;
        o->sun_xz[1] = s[2] / lxz;
    }
    return VG_OK;
}

int vg_gi_default_palette(float *p) {
    if (!p) return fail(VG_E_INVALID, "null palette");
    std::memset(p, 0, 256 * 4 * sizeof(float));
    // default_palette (src/renderer.rs:201-245), material ids src/voxel.rs:100-157
    struct E {
        int m;
        float r, g, b;
    };
    static const E tab[] = {
        {1, 0.94f, 0.85f, 0.55f},  {2, 0.30f, 0.65f, 0.20f},  {3, 0.40f, 0.27f, 0.15f},
        {4, 0.55f, 0.55f, 0.55f},  {13, 0.42f, 0.27f, 0.13f}, {14, 0.30f, 0.58f, 0.20f},
        {15, 0.95f, 0.96f, 0.99f}, {16, 1.40f, 0.40f, 0.10f}, {17, 0.70f, 0.85f, 0.95f},
        {18, 0.90f, 0.96f, 0.99f}, {19, 0.10f, 0.10f, 0.12f}, {20, 0.62f, 0.55f, 0.45f},
        {21, 1.20f, 0.90f, 0.20f}, {22, 0.55f, 0.92f, 0.98f}, {23, 0.85f, 0.82f, 0.72f},
        {24, 0.25f, 0.15f, 0.08f}, {25, 0.55f, 0.78f, 0.34f}, {26, 0.14f, 0.38f, 0.16f},
        {27, 0.90f, 0.42f, 0.15f}, {28, 0.65f, 0.65f, 0.70f}, {29, 1.60f, 0.60f, 0.10f},
        {30, 1.0f, 1.0f, 1.0f},    {31, 0.38f, 0.70f, 0.25f}, {32, 0.28f, 0.50f, 0.30f},
        {33, 0.30f, 0.58f, 0.20f}, {34, 0.78f, 0.68f, 0.38f}, {35, 0.34f, 0.66f, 0.22f},
        {36, 0.28f, 0.55f, 0.19f}, {37, 1.0f, 1.0f, 1.0f},
    };
    for (const E &e : tab) {
        p[e.m * 4 + 0] = e.r;
        p[e.m * 4 + 1] = e.g;
        p[e.m * 4 + 2] = e.b;
        p[e.m * 4 + 3] = 1.0f;
    }
    for (int m = 5; m <= 12; ++m) { // water levels L1..L8
        p[m * 4 + 0] = 0.10f;
        p[m * 4 + 1] = 0.45f;
        p[m * 4 + 2] = 0.75f;
        p[m * 4 + 3] = 1.0f;
    }
    return VG_OK;
}

int vg_gi_directions(int32_t n, float *o) {
    if (n <= 0 || !o) return fail(VG_E_INVALID, "bad direction count %d", n);
    const double golden = M_PI * (3.0 - std::sqrt(5.0)); // 2.39996322972865...
    for (int32_t i = 0; i < n; ++i) {
        const double z = 1.0 - 2.0 * ((double)i + 0.5) / (double)n;
        const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double phi = golden * (double)i;
        o[4 * i + 0] = (float)(r * std::cos(phi));
        o[4 * i + 1] = (float)z;
        o[4 * i + 2] = (float)(r * std::sin(phi));
        o[4 * i + 3] = 0.0f;
    }
    return VG_OK;
}

int vg_gi_bounce_table(float *o) {
    if (!o) return fail(VG_E_INVALID, "null table");
    const int n = VG_GI_BOUNCE_TABLE;
    for (int i = 0; i < n; ++i) {
        const double u1 = ((double)i + 0.5) / (double)n;
        const double u2 = radical_inverse2((uint32_t)i);
        const double r = std::sqrt(u1), phi = 2.0 * M_PI * u2;
        o[4 * i + 0] = (float)(r * std::cos(phi));
        o[4 * i + 1] = (float)(r * std::sin(phi));
        o[4 * i + 2] = (float)std::sqrt(1.0 - u1);
        o[4 * i + 3] = 0.0f;
    }
    return VG_OK;
}

} // extern "C"
