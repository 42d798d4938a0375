/*
 * CPU ORACLE — test infrastructure only (see oracle.h). FP64 restatement of the GI probe
 * update, cs_gi_probe_update / probe_ray_radiance (shaders/gi_probes.wgsl:187-396), with
 * the build's 2-bounce extension.
 *
 * What must equal the device bit for bit is computed in fp32 with the device's operation
 * order: probe origins and relocation offsets (gi_probes.wgsl:298-306, 381-392), gather /
 * shadow / bounce ray origins and directions, the traversal (R1) and the sun-disc test.
 * Everything that is only summed is double: radiance, SH projection (4 pi / N weights),
 * staging sums and the hysteresis fold.
 *
 * Parity pins (tests/test_gi_oracle.py): closed forms (constant sky -> c0 = 4 pi L 0.282095
 * and zero linear terms on the symmetric set; a sealed box -> albedo * ambient * tail;
 * 2-bounce sealed box -> albedo^2 * ambient * tail); sky_color at directions where the
 * reference formula reduces to its constants; the fold/relocation recurrences against a
 * hand-stepped sequence; Fibonacci and bounce tables against their closed forms.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double smooth_d(double e0, double e1, double x) {
    double t = (x - e0) / (e1 - e0);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return t * t * (3.0 - 2.0 * t);
}

void orc_gi_sky(const orc_gi_params *p, const float d[3], double out[3]) {
    const double dy = d[1];
    const double up = dy < -0.2 ? -0.2 : (dy > 1.0 ? 1.0 : dy);
    const double zf = smooth_d(0.0, 0.55, up);
    for (int k = 0; k < 3; ++k)
        out[k] = (double)p->sky_horizon[k] * (1.0 - zf) + (double)p->sky_zenith[k] * zf;
    if (p->dusk_w > 0.0f) {
        const double l = sqrt((double)d[0] * d[0] + (double)d[2] * d[2]);
        double azi = 0.0;
        if (l > 0.0) {
            azi = (d[0] / l) * p->sun_xz[0] + (d[2] / l) * p->sun_xz[1];
            azi = azi < 0.0 ? 0.0 : (azi > 1.0 ? 1.0 : azi);
        }
        const double wgt = (double)p->dusk_w * azi * azi * azi * (1.0 - zf);
        out[0] += 1.50 * wgt;
        out[1] += 0.55 * wgt;
        out[2] += 0.18 * wgt;
    }
    double sa = (double)d[0] * p->sun[0] + (double)d[1] * p->sun[1] + (double)d[2] * p->sun[2];
    sa = sa > 0.0 ? sa : 0.0;
    const double h = pow(sa, 6.0) * (double)p->halo;
    out[0] += 1.0 * h;
    out[1] += 0.78 * h;
    out[2] += 0.52 * h;
    /* the disc is a threshold: decided in fp32 exactly as the device does */
    float saf = d[0] * p->sun[0] + d[1] * p->sun[1] + d[2] * p->sun[2];
    saf = saf > 0.0f ? saf : 0.0f;
    if (saf > 0.9985f) {
        out[0] = out[0] * 0.1 + 2.4 * 0.9;
        out[1] = out[1] * 0.1 + 2.0 * 0.9;
        out[2] = out[2] * 0.1 + 1.5 * 0.9;
    }
}

static uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

uint32_t orc_gi_bounce_index(uint32_t seed, int32_t round, uint32_t probe, uint32_t k) {
    uint32_t x = mix32(seed ^ ((uint32_t)round * 0x85EBCA6Bu));
    x = mix32(x ^ probe);
    x = mix32(x ^ k);
    return x >> 20;
}

static float face_n(uint32_t f, int k) {
    return (f < 6u && (int)(f >> 1) == k) ? ((f & 1u) ? -1.0f : 1.0f) : 0.0f;
}

static float face_ndl(uint32_t f, const float s[3]) {
    if (f >= 6u) return 0.0f;
    const float v = (f & 1u) ? -s[f >> 1] : s[f >> 1];
    return v > 0.0f ? v : 0.0f;
}

static void offset_origin(const float o[3], const float d[3], float t, uint32_t f, float out[3]) {
    for (int k = 0; k < 3; ++k) {
        float p = o[k] + d[k] * t;
        out[k] = p + face_n(f, k) * 0.03f;
    }
}

static void bounce_dir(const float *e, uint32_t f, float out[3]) {
    const int ax = (int)(f >> 1);
    const float sg = (f & 1u) ? -1.0f : 1.0f;
    out[ax] = sg * e[2];
    out[ax == 2 ? 0 : ax + 1] = e[0];
    out[ax == 0 ? 2 : ax - 1] = e[1];
}

/* 1 when the sun is visible from o (no solid voxel within shadow_tmax, water included). */
static double sun_visible(const orc_world *w, const orc_gi_params *p, const float o[3],
                          uint64_t rays[3]) {
    orc_hit h;
    if (orc_trace_dir(w, o, p->sun, p->shadow_tmax, ORC_ANY_HIT, &h, NULL)) rays[1]++;
    return h.hit ? 0.0 : 1.0;
}

/* Radiance arriving at a surface point from its cosine-sampled bounce ray (the tail of a
 * 2-bounce gather sample, and the learned update's training target): so is the offset
 * secondary origin, fX the face the bounce leaves through, e the bounce-table entry. */
static void bounce_tail(const orc_world *w, const orc_gi_params *p, const float *palette,
                        const float so[3], uint32_t fX, const float *e, double lb[3],
                        uint64_t rays[3]) {
    float bd[3];
    bounce_dir(e, fX, bd);
    orc_hit hy;
    if (orc_trace_dir(w, so, bd, p->gi_dist, ORC_SKIP_TRANSPARENT, &hy, NULL)) rays[2]++;
    if (!hy.hit) {
        double sky[3];
        orc_gi_sky(p, bd, sky);
        for (int c = 0; c < 3; ++c) lb[c] = sky[c] * (double)p->gi_sky;
        return;
    }
    const float *aY = palette + 4 * hy.material;
    const double ndlY = face_ndl(hy.face, p->sun);
    float yo[3];
    offset_origin(so, bd, hy.t, hy.face, yo);
    const double visY = ndlY > 0.0 ? sun_visible(w, p, yo, rays) : 0.0;
    for (int c = 0; c < 3; ++c)
        lb[c] = (double)aY[c] * ((double)p->sun_color[c] * ndlY * visY +
                                 (double)p->ambient[c] * (double)p->ambient_tail);
}

void orc_gi_bounce_tail(const orc_world *w, const orc_gi_params *p, const float *btab,
                        const float *palette, const orc_nrc_query *q, uint32_t bidx,
                        double out[3], uint64_t rays[3]) {
    const uint32_t f = q->face & 0xFFu;
    float so[3];
    for (int k = 0; k < 3; ++k) so[k] = q->pos[k] + face_n(f, k) * 0.03f;
    bounce_tail(w, p, palette, so, f, btab + 4 * (bidx & (ORC_GI_BOUNCE_TABLE - 1)), out, rays);
}

static void gi_sample(const orc_world *w, const orc_gi_params *p, const float *btab,
                      const float *palette, const float o[3], const float d[3], uint32_t probe,
                      uint32_t k, const orc_nrc_eval *cache, double out[4], uint8_t *face,
                      float *t_hit, uint64_t rays[3]) {
    orc_hit h;
    if (orc_trace_dir(w, o, d, p->gi_dist, ORC_SKIP_TRANSPARENT, &h, NULL)) rays[0]++;
    if (!h.hit) {
        double sky[3];
        orc_gi_sky(p, d, sky);
        for (int c = 0; c < 3; ++c) out[c] = sky[c] * (double)p->gi_sky;
        out[3] = p->gi_dist;
        *face = ORC_FACE_NONE;
        *t_hit = p->gi_dist;
        return;
    }
    const uint32_t fX = h.face;
    const float tX = h.t;
    const float *aX = palette + 4 * h.material;
    const double ndlX = face_ndl(fX, p->sun);
    float so[3];
    offset_origin(o, d, tX, fX, so);
    const double visX = ndlX > 0.0 ? sun_visible(w, p, so, rays) : 0.0;
    double lb[3];
    if (cache) {
        /* learned: the path ends at the gather hit; the cache gives its bounce tail */
        orc_nrc_query q;
        for (int c = 0; c < 3; ++c) q.pos[c] = o[c] + d[c] * tX;
        for (int c = 0; c < 3; ++c) q.dir[c] = d[c];
        q.face = fX;
        q.aux = 0;
        double y[3];
        orc_nrc_forward(cache->cfg, cache->mlp, cache->grid, 1, &q, cache->emulate, y);
        for (int c = 0; c < 3; ++c) lb[c] = y[c] > 0.0 ? y[c] : 0.0;
    } else if (p->bounces == 2) {
        bounce_tail(w, p, palette, so, fX,
                    btab + 4 * orc_gi_bounce_index(p->seed, p->round, probe, k), lb, rays);
    } else {
        for (int c = 0; c < 3; ++c) lb[c] = (double)p->ambient[c] * (double)p->ambient_tail;
    }
    for (int c = 0; c < 3; ++c)
        out[c] = (double)aX[c] * ((double)p->sun_color[c] * ndlX * visX + lb[c]);
    out[3] = tX;
    *face = (uint8_t)fX;
    *t_hit = tX;
}

void orc_gi_sample(const orc_world *w, const orc_gi_params *p, const float *btab,
                   const float *palette, const float o[3], const float d[3], uint32_t probe,
                   uint32_t k, double out[4], uint8_t *face, float *t_hit, uint64_t rays[3]) {
    gi_sample(w, p, btab, palette, o, d, probe, k, NULL, out, face, t_hit, rays);
}

static int pos_mod_i(int32_t a, int32_t b) {
    int32_t r = a % b;
    return r < 0 ? r + b : r;
}

static void probe_cell(const orc_gi_params *p, const orc_world *w, int64_t idx, int32_t g[3]) {
    const int32_t sx = (int32_t)(idx % p->grid[0]);
    const int32_t sy = (int32_t)((idx / p->grid[0]) % p->grid[1]);
    const int32_t sz = (int32_t)(idx / ((int64_t)p->grid[0] * p->grid[1]));
    const int32_t ox = w->origin[0] / p->spacing, oz = w->origin[2] / p->spacing;
    g[0] = ox + pos_mod_i(sx - ox, p->grid[0]);
    g[1] = sy;
    g[2] = oz + pos_mod_i(sz - oz, p->grid[2]);
}

static int solid_for_gi(uint32_t m) {
    if (m == 0u || m == 33u) return 0;
    if ((m >= 5u && m <= 12u) || m == 18u) return 0;
    return 1;
}

/* SH-L1 basis (gi_probes.wgsl:72-74) in double, the reference's truncated constants. */
static void sh_basis(const float d[3], double b[4]) {
    b[0] = 0.282095;
    b[1] = 0.488603 * d[1];
    b[2] = 0.488603 * d[2];
    b[3] = 0.488603 * d[0];
}

typedef struct {
    const orc_world *w;
    const orc_gi_params *p;
    const float *dirs, *btab, *pal;
    double *probes;
    uint32_t n_upd, phase, epoch;
    const orc_nrc_eval *cache;
    int64_t next;
    pthread_mutex_t mu;
    uint64_t rays[3];
} upd_job;

static void update_probe(upd_job *J, uint32_t j, uint64_t rays[3]) {
    const orc_gi_params *p = J->p;
    const int64_t idx = (int64_t)j * p->update_div + J->phase;
    int32_t g[3];
    probe_cell(p, J->w, idx, g);
    double *pr = J->probes + idx * ORC_PROBE_FIELDS;
    double *info = pr + 20, *st = pr + 24, *st_info = pr + 44;
    const int is_new = info[3] < 0.5 || info[0] != (double)g[0] || info[1] != (double)g[1] ||
                       info[2] != (double)g[2];
    float off[3];
    for (int k = 0; k < 3; ++k) off[k] = is_new ? 0.0f : (float)st_info[1 + k];
    float o[3];
    for (int k = 0; k < 3; ++k) o[k] = ((float)g[k] + 0.5f) * (float)p->spacing + off[k];
    if (p->flags & ORC_GI_SKIP_BURIED) {
        const uint32_t m = orc_material_at_world(J->w, (int32_t)floorf(o[0]), (int32_t)floorf(o[1]),
                                                 (int32_t)floorf(o[2]));
        if (solid_for_gi(m)) {
            memset(pr, 0, ORC_PROBE_FIELDS * sizeof(double));
            info[0] = g[0];
            info[1] = g[1];
            info[2] = g[2];
            return;
        }
    }
    const uint32_t N = (uint32_t)p->rays_per_probe;
    const double wpp = 4.0 * M_PI / (double)N;
    double c[5][4];
    memset(c, 0, sizeof(c));
    float near_t = 1.0e30f;
    uint32_t near_f = ORC_FACE_NONE;
    for (uint32_t k = 0; k < N; ++k) {
        const float *dd = J->dirs + 4 * (k * (uint32_t)p->epochs + J->epoch);
        const float d[3] = {dd[0], dd[1], dd[2]};
        double rad[4];
        uint8_t f;
        float t;
        gi_sample(J->w, p, J->btab, J->pal, o, d, (uint32_t)idx, k, J->cache, rad, &f, &t, rays);
        double b[4];
        sh_basis(d, b);
        const double ws[5] = {rad[0], rad[1], rad[2], rad[3], rad[3] * rad[3]};
        for (int a = 0; a < 5; ++a)
            for (int q = 0; q < 4; ++q) c[a][q] += wpp * ws[a] * b[q];
        if (f != ORC_FACE_NONE && t < near_t) {
            near_t = t;
            near_f = f;
        }
    }
    if (is_new) {
        for (int a = 0; a < 5; ++a)
            for (int q = 0; q < 4; ++q) pr[a * 4 + q] = c[a][q];
        memset(st, 0, 24 * sizeof(double));
    } else {
        for (int a = 0; a < 5; ++a)
            for (int q = 0; q < 4; ++q) st[a * 4 + q] += c[a][q];
        st_info[0] += 1.0;
        if (J->epoch == (uint32_t)p->epochs - 1u) {
            const double cnt = st_info[0] > 1.0 ? st_info[0] : 1.0;
            const double h = p->hysteresis;
            for (int i = 0; i < 20; ++i) pr[i] = (st[i] / cnt) * (1.0 - h) + pr[i] * h;
            memset(st, 0, 24 * sizeof(double));
        }
    }
    /* relocation in fp32, as the device computes it */
    float n[3];
    for (int k = 0; k < 3; ++k) n[k] = face_n(near_f, k);
    if (near_t < 1.2f) {
        const float m = (1.2f - near_t) < 0.35f ? (1.2f - near_t) : 0.35f;
        for (int k = 0; k < 3; ++k) {
            float v = off[k] + n[k] * m;
            off[k] = v < -3.5f ? -3.5f : (v > 3.5f ? 3.5f : v);
        }
    } else if (near_t > 2.5f) {
        for (int k = 0; k < 3; ++k) off[k] = off[k] * 0.98f;
    }
    for (int k = 0; k < 3; ++k) st_info[1 + k] = off[k];
    info[0] = g[0];
    info[1] = g[1];
    info[2] = g[2];
    info[3] = 1.0;
}

static void *upd_worker(void *arg) {
    upd_job *J = (upd_job *)arg;
    uint64_t rays[3] = {0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int64_t j = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (j >= (int64_t)J->n_upd) break;
        update_probe(J, (uint32_t)j, rays);
    }
    pthread_mutex_lock(&J->mu);
    for (int k = 0; k < 3; ++k) J->rays[k] += rays[k];
    pthread_mutex_unlock(&J->mu);
    return NULL;
}

static int clamp_threads(int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    return threads;
}

static void gi_update(const orc_world *w, const orc_gi_params *p, const float *dirs,
                      const float *btab, const float *palette, const orc_nrc_eval *cache,
                      double *probes, int threads, uint64_t rays[3]) {
    const int64_t total = (int64_t)p->grid[0] * p->grid[1] * p->grid[2];
    const uint32_t div = (uint32_t)p->update_div;
    upd_job J;
    memset(&J, 0, sizeof(J));
    J.w = w;
    J.p = p;
    J.dirs = dirs;
    J.btab = btab;
    J.pal = palette;
    J.probes = probes;
    J.cache = cache;
    J.phase = (uint32_t)p->round % div;
    J.epoch = ((uint32_t)p->round / div) % (uint32_t)p->epochs;
    J.n_upd = total > (int64_t)J.phase ? (uint32_t)((total - J.phase + div - 1) / div) : 0u;
    pthread_mutex_init(&J.mu, NULL);
    threads = clamp_threads(threads);
    pthread_t th[256];
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, upd_worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&J.mu);
    if (rays)
        for (int k = 0; k < 3; ++k) rays[k] += J.rays[k];
}

void orc_gi_update(const orc_world *w, const orc_gi_params *p, const float *dirs,
                   const float *btab, const float *palette, double *probes, int threads,
                   uint64_t rays[3]) {
    gi_update(w, p, dirs, btab, palette, NULL, probes, threads, rays);
}

void orc_gi_update_learned(const orc_world *w, const orc_gi_params *p, const float *dirs,
                           const float *btab, const float *palette, const orc_nrc_eval *cache,
                           double *probes, int threads, uint64_t rays[3]) {
    gi_update(w, p, dirs, btab, palette, cache, probes, threads, rays);
}

typedef struct {
    const orc_world *w;
    const orc_gi_params *p;
    int32_t n_dirs;
    const float *dirs, *btab, *pal, *origins;
    const int64_t *slots;
    int64_t n, next;
    double *out;
    pthread_mutex_t mu;
} ref_job;

static void *ref_worker(void *arg) {
    ref_job *J = (ref_job *)arg;
    uint64_t rays[3] = {0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int64_t i = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (i >= J->n) break;
        const float *o = J->origins + 3 * i;
        const double wpp = 4.0 * M_PI / (double)J->n_dirs;
        double c[20];
        memset(c, 0, sizeof(c));
        for (int32_t k = 0; k < J->n_dirs; ++k) {
            const float d[3] = {J->dirs[4 * k], J->dirs[4 * k + 1], J->dirs[4 * k + 2]};
            double rad[4], b[4];
            uint8_t f;
            float t;
            orc_gi_sample(J->w, J->p, J->btab, J->pal, o, d, (uint32_t)J->slots[i], (uint32_t)k,
                          rad, &f, &t, rays);
            sh_basis(d, b);
            const double ws[5] = {rad[0], rad[1], rad[2], rad[3], rad[3] * rad[3]};
            for (int a = 0; a < 5; ++a)
                for (int q = 0; q < 4; ++q) c[a * 4 + q] += wpp * ws[a] * b[q];
        }
        memcpy(J->out + 20 * i, c, sizeof(c));
    }
    return NULL;
}

void orc_gi_reference(const orc_world *w, const orc_gi_params *p, int32_t n_dirs,
                      const float *dirs, const float *btab, const float *palette,
                      const int64_t *slots, const float *origins, int64_t n, double *out,
                      int threads) {
    ref_job J;
    memset(&J, 0, sizeof(J));
    J.w = w;
    J.p = p;
    J.n_dirs = n_dirs;
    J.dirs = dirs;
    J.btab = btab;
    J.pal = palette;
    J.origins = origins;
    J.slots = slots;
    J.n = n;
    J.out = out;
    pthread_mutex_init(&J.mu, NULL);
    threads = clamp_threads(threads);
    pthread_t th[256];
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, ref_worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&J.mu);
}

typedef struct {
    const orc_world *w;
    const orc_gi_params *p;
    const float *btab, *pal;
    const orc_nrc_query *q;
    const uint32_t *bidx;
    double *out;
    int64_t n, next;
    pthread_mutex_t mu;
} tail_job;

static void *tail_worker(void *arg) {
    tail_job *J = (tail_job *)arg;
    uint64_t rays[3] = {0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&J->mu);
        const int64_t i0 = J->next;
        J->next += 256;
        pthread_mutex_unlock(&J->mu);
        if (i0 >= J->n) break;
        const int64_t i1 = i0 + 256 < J->n ? i0 + 256 : J->n;
        for (int64_t i = i0; i < i1; ++i)
            orc_gi_bounce_tail(J->w, J->p, J->btab, J->pal, J->q + i, J->bidx[i], J->out + 3 * i,
                               rays);
    }
    return NULL;
}

void orc_gi_bounce_tails(const orc_world *w, const orc_gi_params *p, const float *btab,
                         const float *palette, int64_t n, const orc_nrc_query *q,
                         const uint32_t *bidx, double *out, int threads) {
    tail_job J;
    memset(&J, 0, sizeof(J));
    J.w = w;
    J.p = p;
    J.btab = btab;
    J.pal = palette;
    J.q = q;
    J.bidx = bidx;
    J.out = out;
    J.n = n;
    pthread_mutex_init(&J.mu, NULL);
    threads = clamp_threads(threads);
    pthread_t th[256];
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, tail_worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&J.mu);
}

/* ------------------------------------------------------------- probe sampling --
 * The probe cache's consumer in fp32, operation for operation as the device evaluates it
 * (dot products left to right, no contraction): sh_eval_irradiance / sh_eval_scalar /
 * sample_probes (shaders/gi_probes.wgsl:69-179) and indirect_light in probe mode
 * (shaders/rt_gi.wgsl:128-139; GI_MAX_T 90 at shaders/raymarch.wgsl:3821, GI_STRENGTH 1).
 * A probe record is 48 floats in GiProbe order: sh_r 0, sh_g 4, sh_b 8, sh_dist 12,
 * sh_dist2 16, info 20, staging 24..43, st_info 44. */
static float dot3f(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

void orc_sh_irradiance(const float *rec, const float n[3], float out[3]) {
    const float a1 = (float)(0.488603 * 0.666667);
    const float lv[3] = {n[1], n[2], n[0]};
    for (int c = 0; c < 3; ++c) {
        const float *sh = rec + 4 * c;
        const float l[3] = {sh[1], sh[2], sh[3]};
        const float dc = sh[0] * 0.282095f;
        const float lin = dot3f(l, lv) * a1;
        const float lmax = sqrtf(dot3f(l, l)) * a1;
        const float q = 0.95f * dc / (lmax > 1e-6f ? lmax : 1e-6f);
        const float sc = q < 1.0f ? q : 1.0f;
        const float e = dc + sc * lin;
        out[c] = e > 0.0f ? e : 0.0f;
    }
}

static float sh_scalar_f(const float *c, const float d[3]) {
    return c[0] * 0.282095f + c[1] * (0.488603f * d[1]) + c[2] * (0.488603f * d[2]) +
           c[3] * (0.488603f * d[0]);
}

void orc_sample_probes(const float *probes, const int32_t grid[3], int32_t spacing,
                       const float pw[3], const float n[3], const float v[3], float out[3]) {
    float pq[3], f[3];
    int32_t base[3];
    const float sp = (float)spacing;
    for (int k = 0; k < 3; ++k) {
        pq[k] = (pw[k] + n[k] * 1.0f) - v[k] * 1.0f;
        const float pg = pq[k] / sp - 0.5f;
        const float fb = floorf(pg);
        base[k] = (int32_t)fb;
        const float fl = pg - fb;
        f[k] = fl * fl * (3.0f - 2.0f * fl);
    }
    float acc[3] = {0.0f, 0.0f, 0.0f}, wsum = 0.0f;
    for (int i = 0; i < 8; ++i) {
        const int32_t off[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
        float w = 1.0f;
        for (int k = 0; k < 3; ++k) w *= off[k] ? f[k] : 1.0f - f[k];
        if (w <= 0.0f) continue;
        int32_t g[3], s[3];
        for (int k = 0; k < 3; ++k) {
            g[k] = base[k] + off[k];
            const int32_t m = g[k] % grid[k];
            s[k] = m < 0 ? m + grid[k] : m;
        }
        const float *rec =
            probes + 48 * (s[0] + (int64_t)s[1] * grid[0] + (int64_t)s[2] * grid[0] * grid[1]);
        if (rec[23] < 0.5f || rec[20] != (float)g[0] || rec[21] != (float)g[1] || rec[22] != (float)g[2])
            continue;
        float dp[3], dn[3];
        for (int k = 0; k < 3; ++k) {
            dp[k] = (((float)g[k] + 0.5f) * sp + rec[45 + k]) - pq[k];
            dn[k] = dp[k] + n[k] * 0.001f;
        }
        const float inv = 1.0f / sqrtf(dot3f(dn, dn));
        const float d[3] = {dn[0] * inv, dn[1] * inv, dn[2] * inv};
        float bf = dot3f(d, n);
        bf = bf > 0.0f ? bf : 0.0f;
        w = w * (bf * bf + 0.05f);
        const float dist = sqrtf(dot3f(dp, dp));
        const float tp[3] = {-d[0], -d[1], -d[2]};
        const float mu = sh_scalar_f(rec + 12, tp);
        float delta = dist - (mu + 0.35f);
        delta = delta > 0.0f ? delta : 0.0f;
        if (delta > 0.0f) {
            const float mu2 = sh_scalar_f(rec + 16, tp);
            float var = mu2 - mu * mu;
            var = var > 0.25f ? var : 0.25f;
            float c = var / (var + delta * delta);
            c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
            w = w * c;
        }
        float e[3];
        orc_sh_irradiance(rec, n, e);
        for (int k = 0; k < 3; ++k) acc[k] = acc[k] + e[k] * w;
        wsum = wsum + w;
    }
    for (int k = 0; k < 3; ++k) out[k] = wsum <= 1e-4f ? 0.0f : acc[k] / wsum;
}

static float smooth_f(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

void orc_frame_indirect(const float *probes, const int32_t grid[3], int32_t spacing,
                        const float *cam, int32_t width, int32_t height, const uint8_t *hits,
                        float *out4) {
    for (int64_t pix = 0; pix < (int64_t)width * height; ++pix) {
        const uint8_t *hr = hits + 20 * pix;
        float t;
        memcpy(&t, hr + 12, 4);
        const uint32_t face = hr[16], hit = hr[18];
        float *o = out4 + 4 * pix;
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        if (!hit || !(t < 90.0f)) continue;
        float d[3], pw[3], n[3], e[3];
        orc_primary_dir(cam, (int32_t)(pix % width), (int32_t)(pix / width), d);
        for (int k = 0; k < 3; ++k) {
            pw[k] = cam[k] + d[k] * t;
            n[k] = face_n(face, k);
        }
        orc_sample_probes(probes, grid, spacing, pw, n, d, e);
        const float s = 1.0f * (1.0f - smooth_f(54.0f, 90.0f, t));
        for (int k = 0; k < 3; ++k) o[k] = e[k] * s;
    }
}
