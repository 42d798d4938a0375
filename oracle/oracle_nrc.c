/*
 * CPU ORACLE — test infrastructure only (see oracle.h). FP64 restatement of the radiance
 * cache the learned probe update uses: a multiresolution hash-grid encoding of the
 * position (Mueller et al. 2022, "Instant Neural Graphics Primitives", section 3: per-level
 * corner hashing with the primes 1, 2654435761, 805459861 and trilinear interpolation)
 * feeding a 3x64 ReLU MLP, its L2 loss gradient and the Adam step (Kingma & Ba 2015,
 * algorithm 1). The cache has no counterpart in the reference: its definition is the one
 * in include/vg.h (vg_nrc_config), and this file is the oracle the device kernels are
 * checked against.
 *
 * Everything is double except what decides an index: the per-level grid coordinate
 * p * scale[l] and its floor are fp32, as on the device, so both pick the same corners.
 * With `emulate` set, the values the device holds in bf16 are rounded to bf16
 * (round-to-nearest-even) at the same points: the parameters, every layer's input
 * (encoding, hidden activations) and every backpropagated layer gradient. Sums stay
 * double, so device and oracle then differ only by fp32 accumulation order.
 *
 * Parity pins (tests/test_nrc_oracle.py): central finite differences of the loss against
 * the analytic gradient (every MLP weight class and sampled grid entries), the encoding's
 * partition of unity and corner closed forms, hash closed forms, Adam against a hand-
 * stepped recurrence, and bf16 rounding against its definition on tie cases.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

double orc_bf16_round(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return f; /* inf / nan pass through */
    const uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7FFFu + lsb) & 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return f;
}

static inline double rnd(double v, int emulate) { return emulate ? orc_bf16_round(v) : v; }

uint32_t orc_nrc_hash(int32_t x, int32_t y, int32_t z, int32_t log2_table) {
    const uint32_t h = ((uint32_t)x * 1u) ^ ((uint32_t)y * 2654435761u) ^ ((uint32_t)z * 805459861u);
    return h & ((1u << log2_table) - 1u);
}

void orc_nrc_scales(int32_t levels, float coarse_cell, float fine_cell, float *out) {
    const double b = levels > 1 ? pow((double)coarse_cell / fine_cell, 1.0 / (levels - 1)) : 1.0;
    for (int l = 0; l < levels; ++l) out[l] = (float)(pow(b, l) / (double)coarse_cell);
}

/* Corner indices (into the whole grid array, feature 0) and trilinear weights of level l. */
static void corners(const orc_nrc_config *c, int l, const float pos[3], int64_t idx[8],
                    double wt[8]) {
    int32_t i0[3];
    double fr[3];
    for (int k = 0; k < 3; ++k) {
        const float s = pos[k] * c->scale[l];
        const float fl = floorf(s);
        i0[k] = (int32_t)fl;
        fr[k] = (double)(s - fl);
    }
    const int64_t T = (int64_t)1 << c->log2_table;
    for (int q = 0; q < 8; ++q) {
        const int cx = q & 1, cy = (q >> 1) & 1, cz = (q >> 2) & 1;
        const uint32_t h = orc_nrc_hash(i0[0] + cx, i0[1] + cy, i0[2] + cz, c->log2_table);
        idx[q] = ((int64_t)l * T + h) * 2;
        wt[q] = (cx ? fr[0] : 1.0 - fr[0]) * (cy ? fr[1] : 1.0 - fr[1]) * (cz ? fr[2] : 1.0 - fr[2]);
    }
}

void orc_nrc_encode(const orc_nrc_config *c, const float *grid, const orc_nrc_query *q,
                    int emulate, double x[ORC_NRC_WIDTH]) {
    memset(x, 0, sizeof(double) * ORC_NRC_WIDTH);
    for (int l = 0; l < c->levels; ++l) {
        int64_t idx[8];
        double wt[8];
        corners(c, l, q->pos, idx, wt);
        double f0 = 0.0, f1 = 0.0;
        for (int k = 0; k < 8; ++k) {
            f0 += wt[k] * rnd(grid[idx[k]], emulate);
            f1 += wt[k] * rnd(grid[idx[k] + 1], emulate);
        }
        x[2 * l] = rnd(f0, emulate);
        x[2 * l + 1] = rnd(f1, emulate);
    }
    const uint32_t face = q->face & 0xFFu;
    if (face < 6) x[ORC_NRC_IN_FACE + face] = 1.0;
    for (int k = 0; k < 3; ++k) x[ORC_NRC_IN_DIR + k] = rnd(q->dir[k], emulate);
    x[ORC_NRC_IN_ONE] = 1.0;
}

/* Layer l's weight (out, in) from the logical parameter vector. */
static inline double wgt(const float *mlp, int l, int o, int i, int emulate) {
    return rnd(mlp[l * ORC_NRC_WIDTH * ORC_NRC_WIDTH + o * ORC_NRC_WIDTH + i], emulate);
}

/* Forward of one query; keeps every layer's input (act[0] = encoding, act[1..3] the
 * hidden activations after ReLU, bf16-rounded under emulate) and the output y[3]. */
static void forward1(const orc_nrc_config *c, const float *mlp, const float *grid,
                     const orc_nrc_query *q, int emulate, double act[4][ORC_NRC_WIDTH],
                     double y[3]) {
    orc_nrc_encode(c, grid, q, emulate, act[0]);
    for (int l = 0; l < 3; ++l)
        for (int o = 0; o < ORC_NRC_WIDTH; ++o) {
            double s = 0.0;
            for (int i = 0; i < ORC_NRC_WIDTH; ++i) s += wgt(mlp, l, o, i, emulate) * act[l][i];
            act[l + 1][o] = rnd(s > 0.0 ? s : 0.0, emulate);
        }
    for (int o = 0; o < 3; ++o) {
        double s = 0.0;
        for (int i = 0; i < ORC_NRC_WIDTH; ++i) s += wgt(mlp, 3, o, i, emulate) * act[3][i];
        y[o] = s;
    }
}

void orc_nrc_forward(const orc_nrc_config *c, const float *mlp, const float *grid, int64_t n,
                     const orc_nrc_query *q, int emulate, double *out3) {
    double act[4][ORC_NRC_WIDTH];
    for (int64_t s = 0; s < n; ++s) forward1(c, mlp, grid, q + s, emulate, act, out3 + 3 * s);
}

double orc_nrc_loss(const orc_nrc_config *c, const float *mlp, const float *grid, int64_t n,
                    const orc_nrc_query *q, const float *target4, int emulate) {
    double act[4][ORC_NRC_WIDTH], y[3], loss = 0.0;
    for (int64_t s = 0; s < n; ++s) {
        forward1(c, mlp, grid, q + s, emulate, act, y);
        const float *t = target4 + 4 * s;
        for (int k = 0; k < 3; ++k) loss += (double)t[3] * (y[k] - t[k]) * (y[k] - t[k]);
    }
    return loss / (2.0 * (double)n);
}

/* Loss and its gradient over the MLP (ORC_NRC_MLP_PARAMS doubles) and the grid (dense,
 * levels * 2^log2_table * 2 doubles); both gradient arrays are accumulated into. */
double orc_nrc_loss_grad(const orc_nrc_config *c, const float *mlp, const float *grid, int64_t n,
                         const orc_nrc_query *q, const float *target4, int emulate,
                         double *g_mlp, double *g_grid) {
    double act[4][ORC_NRC_WIDTH], y[3], loss = 0.0;
    double dy[ORC_NRC_WIDTH], dx[ORC_NRC_WIDTH];
    const double inv_n = 1.0 / (double)n;
    for (int64_t s = 0; s < n; ++s) {
        forward1(c, mlp, grid, q + s, emulate, act, y);
        const float *t = target4 + 4 * s;
        /* output layer gradient (rows >= 3 are zero) */
        memset(dy, 0, sizeof(dy));
        for (int k = 0; k < 3; ++k) {
            const double e = y[k] - t[k];
            loss += (double)t[3] * e * e;
            dy[k] = rnd((double)t[3] * e * inv_n, emulate);
        }
        for (int l = 3; l >= 0; --l) {
            const int nout = l == 3 ? 3 : ORC_NRC_WIDTH;
            double *gw = g_mlp + l * ORC_NRC_WIDTH * ORC_NRC_WIDTH;
            for (int o = 0; o < nout; ++o)
                for (int i = 0; i < ORC_NRC_WIDTH; ++i) gw[o * ORC_NRC_WIDTH + i] += dy[o] * act[l][i];
            for (int i = 0; i < ORC_NRC_WIDTH; ++i) {
                double sacc = 0.0;
                for (int o = 0; o < nout; ++o) sacc += wgt(mlp, l, o, i, emulate) * dy[o];
                dx[i] = sacc;
            }
            if (l > 0) {
                /* through the ReLU that produced act[l] */
                for (int i = 0; i < ORC_NRC_WIDTH; ++i)
                    dy[i] = act[l][i] > 0.0 ? rnd(dx[i], emulate) : 0.0;
            }
        }
        /* dx: gradient of the encoding; scatter the grid part to the corners */
        for (int l = 0; l < c->levels; ++l) {
            int64_t idx[8];
            double wt[8];
            corners(c, l, q[s].pos, idx, wt);
            for (int k = 0; k < 8; ++k) {
                g_grid[idx[k]] += wt[k] * dx[2 * l];
                g_grid[idx[k] + 1] += wt[k] * dx[2 * l + 1];
            }
        }
    }
    return loss * 0.5 * inv_n;
}

/* One Adam step at step number `step` (1-based) over n parameters: m, v, p updated in
 * place (double state). sparse: entries whose gradient is exactly zero are left alone. */
void orc_nrc_adam(int64_t n, double *p, double *m, double *v, const double *g, double lr,
                  double beta1, double beta2, double eps, int64_t step, int sparse) {
    const double c1 = 1.0 - pow(beta1, (double)step), c2 = 1.0 - pow(beta2, (double)step);
    for (int64_t i = 0; i < n; ++i) {
        if (sparse && g[i] == 0.0) continue;
        m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
        p[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}
