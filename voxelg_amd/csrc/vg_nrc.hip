// Radiance cache for gfx950: multiresolution hash-grid encoding + 3x64 ReLU MLP, inference
// and training as bf16 MFMA kernels with fp32 accumulation, Adam with fp32 master weights.
//
// MFMA formulation. Every layer computes Y^T = W X^T with v_mfma_f32_16x16x32_bf16: the A
// operand is a 16x32 block of W (rows = outputs), the B operand a 32x16 block of the layer
// input with the 16 columns = 16 samples (one per lane of a 16-lane group), so a wave holds
// 4 column blocks = 64 samples. The C layout (lane c + 16 g holds rows 4g..4g+3 of sample
// c) is fed back as the next layer's B operand without leaving the registers: the k-th
// input of k-step ks held by lane group g, element j, is feature
//     phi(ks, g, j) = 32 ks + 16 (j / 4) + 4 g + j % 4,
// i.e. elements 0..3 are C-block 2ks and 4..7 are C-block 2ks+1 of the previous layer. The
// weight fragments are pre-permuted by phi (k_nrc_adam_mlp packs them after every step)
// and staged in LDS. The input layer uses the same phi over the logical input order, so a
// lane encodes hash levels {2g, 2g+1, 8+2g, 9+2g} of its sample.
// Training stages a layer's input and output gradient (bf16) through LDS to form the
// weight gradient dW = dY X^T as MFMAs over samples; per-block partial sums are reduced in
// a second kernel (deterministic). The grid gradient: every (sample, level, corner)
// contribution becomes a record that a counting pass (per-block bucket counts, one exclusive
// scan, vg_scan.h) places in its bucket of 4096 entries; one block per bucket sums its records in LDS
// in 64-bit fixed point, which is exact and therefore independent of the record order, and
// runs Adam on exactly the entries that received a gradient.
#include "vg_nrc.h"

#include "vg_scan.h"

#include <vector>

namespace vg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

struct NrcDev {
    const uint32_t *grid_h;
    const u32x4_t *frag_fwd;
    const u32x4_t *frag_bwd;
    float scale[VG_NRC_MAX_LEVELS];
    int32_t levels, log2_table;
};

constexpr int kSRow = kNrcTile + 8; // bf16 per LDS staging row (padded)

// ----------------------------------------------------------------- helpers --
__device__ __forceinline__ uint32_t bf16_bits(float f) { // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    return bf16_bits(a) | (bf16_bits(b) << 16);
}
__device__ __forceinline__ float bf_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf_hi(uint32_t v) { return __uint_as_float(v & 0xFFFF0000u); }
__device__ __forceinline__ float relu(float v) { return v > 0.0f ? v : 0.0f; }
// running max of |x|, +inf once any x is not finite
__device__ __forceinline__ float gmax_acc(float g, float x) {
    const float a = fabsf(x);
    return a <= 3.402823466e38f ? fmaxf(g, a) : INFINITY;
}

__device__ __forceinline__ f32x4_t mfma(u32x4_t a, u32x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                   __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

// Grid entries (level base included) and trilinear weights of the 8 corners of level l.
__device__ __forceinline__ void level_corners(const NrcDev &m, int l, const float p[3],
                                              uint32_t idx[8], float w[8]) {
    const float s = m.scale[l];
    float fr[3];
    int32_t i0[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float v = p[k] * s;
        const float fl = floorf(v);
        i0[k] = (int32_t)fl;
        fr[k] = v - fl;
    }
    const uint32_t mask = (1u << m.log2_table) - 1u, base = (uint32_t)l << m.log2_table;
    const uint32_t hx[2] = {(uint32_t)i0[0], (uint32_t)i0[0] + 1u};
    const uint32_t hy[2] = {(uint32_t)i0[1] * 2654435761u, ((uint32_t)i0[1] + 1u) * 2654435761u};
    const uint32_t hz[2] = {(uint32_t)i0[2] * 805459861u, ((uint32_t)i0[2] + 1u) * 805459861u};
    const float wx[2] = {1.0f - fr[0], fr[0]}, wy[2] = {1.0f - fr[1], fr[1]},
                wz[2] = {1.0f - fr[2], fr[2]};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int cx = q & 1, cy = (q >> 1) & 1, cz = q >> 2;
        idx[q] = base + ((hx[cx] ^ hy[cy] ^ hz[cz]) & mask);
        w[q] = (wx[cx] * wy[cy]) * wz[cz];
    }
}

__device__ __forceinline__ void level_feat(const NrcDev &m, int l, const float p[3], float &f0,
                                           float &f1) {
    f0 = 0.0f;
    f1 = 0.0f;
    if (l >= m.levels) return;
    uint32_t idx[8];
    float w[8];
    level_corners(m, l, p, idx, w);
    uint32_t v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = m.grid_h[idx[q]];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        f0 = __builtin_fmaf(w[q], bf_lo(v[q]), f0);
        f1 = __builtin_fmaf(w[q], bf_hi(v[q]), f1);
    }
}

// Second input k-step of lane group g: inputs 32 + 4g + j (face one-hot, direction, the
// constant 1; elements 4..7 are inputs >= 48, all zero).
__device__ __forceinline__ u32x4_t input_extras(int g, const float d[3], uint32_t face) {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * g + j; // input 32 + k
        float x = 0.0f;
        if (k < 6) x = (face == (uint32_t)k) ? 1.0f : 0.0f;
        else if (k < 9) x = d[k - 6];
        else if (k == 9) x = 1.0f;
        v[j] = x;
    }
    return u32x4_t{pack2(v[0], v[1]), pack2(v[2], v[3]), 0u, 0u};
}

// Input k-steps of lane group g for one sample (phi order over the logical inputs): the
// first holds hash levels 2g, 2g+1, 8+2g, 9+2g as bf16 feature pairs.
// The 32 gathers of the four levels are issued together and consumed once (an empty asm
// statement keeps the compiler from waiting between levels); a level beyond the config
// reads entry 0 with zero weights, which gives exactly the zero feature.
__device__ __forceinline__ void build_input(const NrcDev &m, int g, const float p[3],
                                            const float d[3], uint32_t face, u32x4_t &x0,
                                            u32x4_t &x1) {
    const int ls[4] = {2 * g, 2 * g + 1, 8 + 2 * g, 9 + 2 * g};
    uint32_t v[4][8];
    float w[4][8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t idx[8];
        const bool used = ls[k] < m.levels;
        level_corners(m, used ? ls[k] : 0, p, idx, w[k]);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (!used) w[k][q] = 0.0f;
            v[k][q] = m.grid_h[used ? idx[q] : 0u];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        asm volatile("" ::"v"(v[k][0]), "v"(v[k][1]), "v"(v[k][2]), "v"(v[k][3]), "v"(v[k][4]),
                     "v"(v[k][5]), "v"(v[k][6]), "v"(v[k][7]));
    float f[4][2];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[k][0] = 0.0f;
        f[k][1] = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            f[k][0] = __builtin_fmaf(w[k][q], bf_lo(v[k][q]), f[k][0]);
            f[k][1] = __builtin_fmaf(w[k][q], bf_hi(v[k][q]), f[k][1]);
        }
    }
    x0 = u32x4_t{pack2(f[0][0], f[0][1]), pack2(f[1][0], f[1][1]), pack2(f[2][0], f[2][1]),
                 pack2(f[3][0], f[3][1])};
    x1 = input_extras(g, d, face);
}

__device__ __forceinline__ void load_query(const vg_nrc_query *q, uint32_t s, bool valid,
                                           float p[3], float d[3], uint32_t &face, uint32_t &aux) {
    if (valid) {
        const float4 *r = reinterpret_cast<const float4 *>(q + s);
        const float4 a = r[0], b = r[1];
        p[0] = a.x;
        p[1] = a.y;
        p[2] = a.z;
        face = __float_as_uint(a.w);
        d[0] = b.x;
        d[1] = b.y;
        d[2] = b.z;
        aux = __float_as_uint(b.w);
    } else {
        p[0] = p[1] = p[2] = 0.0f;
        d[0] = d[1] = d[2] = 0.0f;
        face = VG_FACE_NONE;
        aux = 0u;
    }
}

__device__ __forceinline__ void load_frags(u32x4_t *dst, const u32x4_t *src) {
    for (int i = threadIdx.x; i < kNrcFrags * 64; i += blockDim.x) dst[i] = src[i];
}

// acc[ob][sb] = sum over k-steps of fragment (lyr, ob, ks) x input[ks][sb], for the wave's
// SB column blocks of 16 samples.
template <int NOB, int SB = 4>
__device__ __forceinline__ void mlp_layer(const u32x4_t *sf, int lyr, const u32x4_t x[2][SB],
                                          f32x4_t acc[4][SB], int lane) {
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob) {
#pragma unroll
        for (int sb = 0; sb < SB; ++sb) acc[ob][sb] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const u32x4_t a = sf[((lyr * 4 + ob) * 2 + ks) * 64 + lane];
#pragma unroll
            for (int sb = 0; sb < SB; ++sb) acc[ob][sb] = mfma(a, x[ks][sb], acc[ob][sb]);
        }
    }
}

template <int SB = 4>
__device__ __forceinline__ void relu_pack(const f32x4_t acc[4][SB], u32x4_t x[2][SB]) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int sb = 0; sb < SB; ++sb) {
            const f32x4_t a = acc[2 * ks][sb], b = acc[2 * ks + 1][sb];
            x[ks][sb] = u32x4_t{pack2(relu(a[0]), relu(a[1])), pack2(relu(a[2]), relu(a[3])),
                                pack2(relu(b[0]), relu(b[1])), pack2(relu(b[2]), relu(b[3]))};
        }
}

// ---------------------------------------------------------------- inference --
#ifndef VG_NRC_INFER_SB
#define VG_NRC_INFER_SB 1
#endif
constexpr int kInferSB = VG_NRC_INFER_SB;
// A wave evaluates SB column blocks of 16 queries (16 SB per tile): small tiles give the
// gathers of a launch of a few ten thousand queries enough waves to overlap.
template <bool PROBE, int SB>
__global__ __launch_bounds__(256) void k_nrc_infer(NrcDev m, const vg_nrc_query *__restrict__ qb,
                                                   size_t q_stride, const uint32_t *sel,
                                                   const uint32_t *n_dev, uint32_t n_max,
                                                   float4 *__restrict__ out,
                                                   float4 *__restrict__ rad,
                                                   const float4 *__restrict__ pal) {
    __shared__ u32x4_t sf[kNrcFrags * 64];
    const vg_nrc_query *q = qb + (sel ? (size_t)sel[0] * q_stride : (size_t)0);
    uint32_t n = n_dev ? *n_dev : n_max;
    n = n < n_max ? n : n_max;
    // the grid is sized for n_max; a block with no tile leaves before staging the weights
    // (the whole block together, so no barrier is left waiting)
    if ((uint64_t)blockIdx.x * 4u * (16u * SB) >= n) return;
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const uint32_t wave = threadIdx.x >> 6;
    constexpr uint32_t kTile = 16u * SB;
    const uint32_t stride = gridDim.x * 4u;
    float p[SB][3], d[SB][3];
    uint32_t face[SB], aux[SB];
    auto load_tile = [&](uint32_t t) {
#pragma unroll
        for (int sb = 0; sb < SB; ++sb) {
            const uint32_t s = t * kTile + (uint32_t)(sb * 16 + c);
            load_query(q, s, s < n, p[sb], d[sb], face[sb], aux[sb]);
        }
    };
    // the wave's first queries load together with the weight staging
    uint32_t tile = blockIdx.x * 4u + wave;
    if ((uint64_t)tile * kTile < n) load_tile(tile);
    load_frags(sf, m.frag_fwd);
    __syncthreads();
    for (; (uint64_t)tile * kTile < n; tile += stride) {
        // probe epilogue: the sample's radiance and the hit's albedo load with the gathers
        float4 r0[SB], a0[SB];
#pragma unroll
        for (int sb = 0; sb < SB; ++sb) {
            r0[sb] = a0[sb] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const uint32_t s = tile * kTile + (uint32_t)(sb * 16 + c);
            if (PROBE && g == 0 && s < n) {
                a0[sb] = pal[(face[sb] >> 8) & 0xFFu];
                r0[sb] = rad[aux[sb]];
            }
        }
        u32x4_t x[2][SB];
#pragma unroll
        for (int sb = 0; sb < SB; ++sb)
            build_input(m, g, p[sb], d[sb], face[sb] & 0xFFu, x[0][sb], x[1][sb]);
        f32x4_t acc[4][SB];
        mlp_layer<4, SB>(sf, 0, x, acc, lane);
        relu_pack<SB>(acc, x);
        mlp_layer<4, SB>(sf, 1, x, acc, lane);
        relu_pack<SB>(acc, x);
        mlp_layer<4, SB>(sf, 2, x, acc, lane);
        relu_pack<SB>(acc, x);
        mlp_layer<1, SB>(sf, 3, x, acc, lane);
        if (g == 0) {
#pragma unroll
            for (int sb = 0; sb < SB; ++sb) {
                const uint32_t s = tile * kTile + (uint32_t)(sb * 16 + c);
                if (s >= n) continue;
                const f32x4_t y = acc[0][sb];
                if (PROBE) {
                    float4 r = r0[sb];
                    r.x += a0[sb].x * relu(y[0]);
                    r.y += a0[sb].y * relu(y[1]);
                    r.z += a0[sb].z * relu(y[2]);
                    rad[aux[sb]] = r;
                } else {
                    out[s] = make_float4(y[0], y[1], y[2], 0.0f);
                }
            }
        }
        if ((uint64_t)(tile + stride) * kTile < n) load_tile(tile + stride);
    }
}

// ----------------------------------------------------------------- training --
// x[ks][sb] element j -> LDS row phi(ks, g, j) (rows >= nrows dropped), column = sample.
__device__ __forceinline__ void stage(uint16_t *s, const u32x4_t x[2][4], int nrows, int g, int c,
                                      int wave) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            const int col = wave * 64 + sb * 16 + c;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t u = x[ks][sb][e];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int j = 2 * e + h;
                    const int row = 32 * ks + 16 * (j >> 2) + 4 * g + (j & 3);
                    if (row < nrows) s[row * kSRow + col] = (uint16_t)(h ? (u >> 16) : (u & 0xFFFFu));
                }
            }
        }
}

// dW block (ob, ib) over the tile's 256 samples: A = D rows 16 ob + r, B = X rows 16 ib + r,
// k = samples 32 kk + 8 g + j. Lane c + 16 g gets dW[16 ob + 4 g + i][16 ib + c].
__device__ __forceinline__ f32x4_t dw_block(const uint16_t *sD, const uint16_t *sX, int ob, int ib,
                                            int g, int r) {
    f32x4_t acc = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < kNrcTile / 32; ++kk) {
        const u32x4_t a =
            *reinterpret_cast<const u32x4_t *>(sD + (16 * ob + r) * kSRow + 32 * kk + 8 * g);
        const u32x4_t b =
            *reinterpret_cast<const u32x4_t *>(sX + (16 * ib + r) * kSRow + 32 * kk + 8 * g);
        acc = mfma(a, b, acc);
    }
    return acc;
}

// Weight gradient of layer lyr for this tile, added to the block partial (stored on the
// block's first tile). Every (block, weight) entry belongs to one wave.
__device__ __forceinline__ void weight_grad(uint16_t *sX, uint16_t *sD, const u32x4_t X[2][4],
                                            const u32x4_t D[2][4], int lyr, float *part,
                                            bool first, int g, int c, int wave) {
    const int nrows = lyr == 3 ? 16 : 64;
    stage(sX, X, 64, g, c, wave);
    stage(sD, D, nrows, g, c, wave);
    __syncthreads();
    const int nblk = (nrows / 16) * 4;
    for (int b = wave; b < nblk; b += 4) {
        const int ob = b >> 2, ib = b & 3;
        const f32x4_t a = dw_block(sD, sX, ob, ib, g, c);
        float *dst = part + lyr * 4096 + (16 * ob + 4 * g) * 64 + 16 * ib + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[i * 64] = first ? a[i] : dst[i * 64] + a[i];
    }
    __syncthreads();
}

// D_in = (W^T D_out) through the ReLU of the layer input Xin, packed as the next B operand.
__device__ __forceinline__ void back_layer(const u32x4_t *sbf, int lyr, int nks,
                                           const u32x4_t Dout[2][4], const u32x4_t Xin[2][4],
                                           u32x4_t Din[2][4], int lane) {
    f32x4_t acc[4][4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) acc[ib][sb] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < nks; ++ks) {
            const u32x4_t a = sbf[((lyr * 4 + ib) * 2 + ks) * 64 + lane];
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) acc[ib][sb] = mfma(a, Dout[ks][sb], acc[ib][sb]);
        }
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            uint32_t o[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int j0 = 2 * e, j1 = 2 * e + 1;
                const float v0 = acc[2 * ks + (j0 >> 2)][sb][j0 & 3];
                const float v1 = acc[2 * ks + (j1 >> 2)][sb][j1 & 3];
                const uint32_t u = Xin[ks][sb][e];
                o[e] = pack2((u & 0xFFFFu) ? v0 : 0.0f, (u >> 16) ? v1 : 0.0f);
            }
            Din[ks][sb] = u32x4_t{o[0], o[1], o[2], o[3]};
        }
}

// Encoding of a training batch: one thread per (sample, level) writes the level's bf16
// feature pair, feat[s * 16 + l] (zero for levels beyond the config); clears the step's
// gradient maximum.
__global__ __launch_bounds__(256) void k_nrc_encode(NrcDev m, const vg_nrc_query *__restrict__ q,
                                                    uint32_t n, uint32_t *__restrict__ feat,
                                                    uint32_t *gmax) {
    const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (t == 0) *gmax = 0u; // the step's gradient maximum is gathered after the backward pass
    const uint32_t s = (uint32_t)(t >> 4), l = (uint32_t)(t & 15u);
    if (s >= n) return;
    float f0 = 0.0f, f1 = 0.0f;
    if ((int)l < m.levels) {
        const float4 a = reinterpret_cast<const float4 *>(q + s)[0];
        const float p[3] = {a.x, a.y, a.z};
        level_feat(m, (int)l, p, f0, f1);
    }
    feat[t] = pack2(f0, f1);
}

// ------------------------------------------------------------ grid gradient --
// Blocks of the record count / scatter passes; each takes a contiguous range of
// (sample, level) pairs, pair p = s * levels + l.
constexpr uint32_t kGBlocks = 256;
constexpr uint32_t kGThreads = 1024; // count / scatter block size
constexpr uint32_t kGRThreads = 512; // per-bucket sum block size

__device__ __forceinline__ void pair_range(uint32_t pairs, uint32_t &p0, uint32_t &p1) {
    const uint32_t per = (pairs + kGBlocks - 1) / kGBlocks;
    p0 = blockIdx.x * per;
    p1 = p0 + per < pairs ? p0 + per : pairs;
}

// Record counts per (bucket, block): hist[bucket * kGBlocks + block]. Pairs whose gradient is
// zero emit no records.
__global__ __launch_bounds__(kGThreads) void k_nrc_gcount(NrcDev m, const vg_nrc_query *__restrict__ q,
                                                    const float2 *__restrict__ dfeat, uint32_t n,
                                                    int gbits, uint32_t nbuckets,
                                                    uint32_t *__restrict__ hist) {
    extern __shared__ uint32_t cnt[];
    for (uint32_t b = threadIdx.x; b < nbuckets; b += kGThreads) cnt[b] = 0u;
    __syncthreads();
    uint32_t p0, p1;
    pair_range(n * (uint32_t)m.levels, p0, p1);
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += kGThreads) {
        const uint32_t s = p / (uint32_t)m.levels, l = p - s * (uint32_t)m.levels;
        const float2 d = dfeat[(size_t)s * 16u + l];
        if (d.x == 0.0f && d.y == 0.0f) continue;
        const float4 a = reinterpret_cast<const float4 *>(q + s)[0];
        const float pos[3] = {a.x, a.y, a.z};
        uint32_t idx[8];
        float w[8];
        level_corners(m, (int)l, pos, idx, w);
#pragma unroll
        for (int k = 0; k < 8; ++k) atomicAdd(&cnt[idx[k] >> gbits], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbuckets; b += kGThreads) hist[(size_t)b * kGBlocks + blockIdx.x] = cnt[b];
}

// The same records, written at their bucket's place: the scanned hist gives every (bucket,
// block) its first slot; inside a block the slots are taken in any order (the sums below
// do not depend on it).
__global__ __launch_bounds__(kGThreads) void k_nrc_gscatter(NrcDev m, const vg_nrc_query *__restrict__ q,
                                                      const float2 *__restrict__ dfeat, uint32_t n,
                                                      int gbits, uint32_t nbuckets,
                                                      const uint32_t *__restrict__ hoff,
                                                      uint32_t *__restrict__ rk,
                                                      float2 *__restrict__ rv) {
    extern __shared__ uint32_t cur[];
    for (uint32_t b = threadIdx.x; b < nbuckets; b += kGThreads) cur[b] = hoff[(size_t)b * kGBlocks + blockIdx.x];
    __syncthreads();
    const uint32_t lmask = (1u << gbits) - 1u;
    uint32_t p0, p1;
    pair_range(n * (uint32_t)m.levels, p0, p1);
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += kGThreads) {
        const uint32_t s = p / (uint32_t)m.levels, l = p - s * (uint32_t)m.levels;
        const float2 d = dfeat[(size_t)s * 16u + l];
        if (d.x == 0.0f && d.y == 0.0f) continue;
        const float4 a = reinterpret_cast<const float4 *>(q + s)[0];
        const float pos[3] = {a.x, a.y, a.z};
        uint32_t idx[8];
        float w[8];
        level_corners(m, (int)l, pos, idx, w);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t slot = atomicAdd(&cur[idx[k] >> gbits], 1u);
            rk[slot] = idx[k] & lmask;
            rv[slot] = make_float2(w[k] * d.x, w[k] * d.y);
        }
    }
}

__device__ __forceinline__ void adam1(float &p, float &m, float &v, float g, float lr, float b1,
                                      float b2, float eps, float c1, float c2) {
    m = b1 * m + (1.0f - b1) * g;
    v = b2 * v + (1.0f - b2) * (g * g);
    p -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// One block per bucket of 2^gbits entries: its records are summed in LDS in 64-bit fixed
// point (scale 2^K with K = 38 - e, |contribution| <= gmax < 2^e: at most 2^23 records per
// entry keep every sum below 2^61; integer sums are exact, so the result does not depend on
// the order of the records). Then, per entry and feature with a non-zero sum: Adam (adam set;
// the bf16 copy the kernels read is refreshed), or the dense gradient (dense set). A feature
// whose sum is zero is left alone, like an entry no sample touched.
__global__ __launch_bounds__(kGRThreads) void k_nrc_greduce(const uint32_t *__restrict__ hoff,
                                                     const uint32_t *__restrict__ rk,
                                                     const float2 *__restrict__ rv, int gbits,
                                                     const uint32_t *gmax, float2 *__restrict__ p,
                                                     float2 *__restrict__ mm, float2 *__restrict__ vv,
                                                     uint32_t *__restrict__ ph,
                                                     float2 *__restrict__ dense, int adam, float lr,
                                                     float b1, float b2, float eps,
                                                     const uint32_t *step) {
    extern __shared__ unsigned long long acc[]; // 2 per entry
    const uint32_t ne = 1u << gbits;
    for (uint32_t i = threadIdx.x; i < 2u * ne; i += kGRThreads) acc[i] = 0ull;
    __syncthreads();
    const uint32_t b = blockIdx.x;
    const uint32_t r0 = hoff[(size_t)b * kGBlocks], r1 = hoff[(size_t)(b + 1u) * kGBlocks];
    const float gm = __uint_as_float(*gmax);
    int e = 0;
    (void)frexpf(gm, &e); // gm < 2^e
    const int K = 38 - e;
    const double up = ldexp(1.0, K), down = ldexp(1.0, -K);
    if (gm > 0.0f && gm <= 3.4e38f) // a non-finite backward pass contributes nothing
        for (uint32_t r = r0 + threadIdx.x; r < r1; r += kGRThreads) {
            const uint32_t k = rk[r];
            const float2 v = rv[r];
            atomicAdd(&acc[2u * k], (unsigned long long)__double2ll_rn((double)v.x * up));
            atomicAdd(&acc[2u * k + 1u], (unsigned long long)__double2ll_rn((double)v.y * up));
        }
    __syncthreads();
    float c1 = 1.0f, c2 = 1.0f;
    if (adam) {
        const float t = (float)*step;
        c1 = 1.0f - powf(b1, t);
        c2 = 1.0f - powf(b2, t);
    }
    for (uint32_t i = threadIdx.x; i < ne; i += kGRThreads) {
        const size_t entry = ((size_t)b << gbits) + i;
        const float gx = (float)((double)(long long)acc[2u * i] * down);
        const float gy = (float)((double)(long long)acc[2u * i + 1u] * down);
        if (dense) dense[entry] = make_float2(gx, gy);
        if (!adam || (gx == 0.0f && gy == 0.0f)) continue;
        float2 pp = p[entry], m1 = mm[entry], v1 = vv[entry];
        if (gx != 0.0f) adam1(pp.x, m1.x, v1.x, gx, lr, b1, b2, eps, c1, c2);
        if (gy != 0.0f) adam1(pp.y, m1.y, v1.y, gy, lr, b1, b2, eps, c1, c2);
        p[entry] = pp;
        mm[entry] = m1;
        vv[entry] = v1;
        ph[entry] = pack2(pp.x, pp.y);
    }
}

// Forward and backward of a training batch from its encoding (feat): weight-gradient sums
// per block into `partial`, the encoding gradient per sample into dfeat[s * 32 + input], and
// the largest |gradient| of the used levels' inputs into gmax (fp32 bits; zeroed by
// k_nrc_encode): the grid gradient's fixed-point scale.
__global__ __launch_bounds__(256) void k_nrc_train(NrcDev m, const vg_nrc_query *__restrict__ q,
                                                   const uint32_t *__restrict__ feat,
                                                   const float4 *__restrict__ tgt, uint32_t n,
                                                   float inv_n, float *__restrict__ partial,
                                                   float *__restrict__ dfeat, uint32_t *gmax) {
    __shared__ u32x4_t sf[kNrcFrags * 64];
    __shared__ u32x4_t sbw[kNrcFrags * 64];
    __shared__ __align__(16) uint16_t sX[kNrcWidth * kSRow];
    __shared__ __align__(16) uint16_t sD[kNrcWidth * kSRow];
    __shared__ float sl[4], sm[4];
    load_frags(sf, m.frag_fwd);
    load_frags(sbw, m.frag_bwd);
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int wave = threadIdx.x >> 6;
    float *part = partial + (size_t)blockIdx.x * kNrcPartial;
    float loss = 0.0f, gmx = 0.0f;
    bool first = true;
    for (uint32_t tile = blockIdx.x; (uint64_t)tile * kNrcTile < n; tile += gridDim.x) {
        u32x4_t X[4][2][4];
        float tw[4], tr[4][3];
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            const uint32_t s = tile * kNrcTile + (uint32_t)(wave * 64 + sb * 16 + c);
            const bool valid = s < n;
            float pos[3], d[3];
            uint32_t face, aux;
            load_query(q, s, valid, pos, d, face, aux);
            const float4 t = valid ? tgt[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            tr[sb][0] = t.x;
            tr[sb][1] = t.y;
            tr[sb][2] = t.z;
            tw[sb] = t.w;
            if (valid) {
                const uint2 lo = *reinterpret_cast<const uint2 *>(feat + (size_t)s * 16u + 2u * g);
                const uint2 hi = *reinterpret_cast<const uint2 *>(feat + (size_t)s * 16u + 8u + 2u * g);
                X[0][0][sb] = u32x4_t{lo.x, lo.y, hi.x, hi.y};
            } else {
                X[0][0][sb] = u32x4_t{0u, 0u, 0u, 0u};
            }
            X[0][1][sb] = input_extras(g, d, face & 0xFFu);
        }
        f32x4_t acc[4][4];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            mlp_layer<4>(sf, l, X[l], acc, lane);
            relu_pack(acc, X[l + 1]);
        }
        mlp_layer<1>(sf, 3, X[3], acc, lane);
        // output gradient: rows 0..2 of lane group 0
        u32x4_t D[2][4], Dn[2][4];
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            float dy[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float e = 0.0f;
                if (g == 0 && i < 3) {
                    const float r = acc[0][sb][i] - tr[sb][i];
                    e = tw[sb] * r;
                    loss += e * r;
                    gmx = fabsf(e) <= 3.402823466e38f ? gmx : INFINITY;
                }
                dy[i] = e * inv_n;
            }
            D[0][sb] = u32x4_t{pack2(dy[0], dy[1]), pack2(dy[2], dy[3]), 0u, 0u};
            D[1][sb] = u32x4_t{0u, 0u, 0u, 0u};
        }
        weight_grad(sX, sD, X[3], D, 3, part, first, g, c, wave);
        back_layer(sbw, 3, 1, D, X[3], Dn, lane);
        weight_grad(sX, sD, X[2], Dn, 2, part, first, g, c, wave);
        back_layer(sbw, 2, 2, Dn, X[2], D, lane);
        weight_grad(sX, sD, X[1], D, 1, part, first, g, c, wave);
        back_layer(sbw, 1, 2, D, X[1], Dn, lane);
        weight_grad(sX, sD, X[0], Dn, 0, part, first, g, c, wave);
        // encoding gradient (inputs 0..31 = C-blocks 0 and 1), scattered to the grid corners
        f32x4_t ga[2][4];
#pragma unroll
        for (int ib = 0; ib < 2; ++ib) {
#pragma unroll
            for (int sb = 0; sb < 4; ++sb) ga[ib][sb] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const u32x4_t a = sbw[((0 * 4 + ib) * 2 + ks) * 64 + lane];
#pragma unroll
                for (int sb = 0; sb < 4; ++sb) ga[ib][sb] = mfma(a, Dn[ks][sb], ga[ib][sb]);
            }
        }
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
            const uint32_t s = tile * kNrcTile + (uint32_t)(wave * 64 + sb * 16 + c);
            if (s >= n) continue;
            float4 *dst = reinterpret_cast<float4 *>(dfeat + (size_t)s * 32u);
            dst[g] = make_float4(ga[0][sb][0], ga[0][sb][1], ga[0][sb][2], ga[0][sb][3]);
            dst[4 + g] = make_float4(ga[1][sb][0], ga[1][sb][1], ga[1][sb][2], ga[1][sb][3]);
            // inputs 4g + k (levels 2g, 2g + 1) and 16 + 4g + k (levels 8 + 2g, 9 + 2g)
            // a non-finite gradient makes the maximum +inf (fmaxf alone would drop a NaN),
            // and a step whose maximum is not finite changes no parameter
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (2 * g + (k >> 1) < m.levels) gmx = gmax_acc(gmx, ga[0][sb][k]);
                if (8 + 2 * g + (k >> 1) < m.levels) gmx = gmax_acc(gmx, ga[1][sb][k]);
            }
        }
        first = false;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        loss += __shfl_xor(loss, off, 64);
        gmx = fmaxf(gmx, __shfl_xor(gmx, off, 64));
    }
    if (lane == 0) {
        sl[wave] = loss;
        sm[wave] = gmx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[VG_NRC_MLP_PARAMS] = ((sl[0] + sl[1]) + (sl[2] + sl[3])) * 0.5f * inv_n;
        atomicMax(gmax, __float_as_uint(fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]))));
    }
}

// Sum of the block partials: the MLP gradient and the loss; with `update`, one more step.
// A block of 1024 threads owns 64 consecutive entries; its 16 thread rows sum every 16th
// block partial, then the rows are added in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void k_nrc_reduce(const float *__restrict__ partial, int nblocks,
                                                     float *mlp_g, float *loss, uint32_t *step,
                                                     int update) {
    __shared__ float sr[16][64];
    const int pl = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + pl;
    float acc = 0.0f;
    if (p <= VG_NRC_MLP_PARAMS)
        for (int b = row; b < nblocks; b += 16) acc += partial[(size_t)b * kNrcPartial + p];
    sr[row][pl] = acc;
    __syncthreads();
    if (row != 0 || p > VG_NRC_MLP_PARAMS) return;
    float sum = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) sum += sr[r][pl];
    if (p < VG_NRC_MLP_PARAMS) {
        mlp_g[p] = sum;
    } else {
        *loss = sum;
        if (update) *step += 1u;
    }
}

// bf16 copy of the whole grid (after an upload or initialisation).
__global__ void k_nrc_grid_pack(const float2 *__restrict__ p, uint32_t *__restrict__ ph,
                                int64_t n_entries) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_entries) return;
    const float2 v = p[e];
    ph[e] = pack2(v.x, v.y);
}

// Dense Adam over the MLP (with `update`), then the forward and backward fragments of the
// (new) weights. One block.
__global__ __launch_bounds__(1024) void k_nrc_adam_mlp(float *__restrict__ p, float *__restrict__ mm,
                                                       float *__restrict__ vv,
                                                       const float *__restrict__ g,
                                                       uint16_t *__restrict__ ffwd,
                                                       uint16_t *__restrict__ fbwd, float lr,
                                                       float b1, float b2, float eps,
                                                       const uint32_t *step, int update) {
    if (update) {
        const float t = (float)*step;
        const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
        for (int i = threadIdx.x; i < VG_NRC_MLP_PARAMS; i += blockDim.x) {
            float pp = p[i], m = mm[i], v = vv[i];
            adam1(pp, m, v, g[i], lr, b1, b2, eps, c1, c2);
            p[i] = pp;
            mm[i] = m;
            vv[i] = v;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kNrcFrags * 64 * 8; e += blockDim.x) {
        const int f = e >> 9, lane = (e >> 3) & 63, j = e & 7;
        const int lyr = f >> 3, blk = (f >> 1) & 3, ks = f & 1;
        const int r = lane & 15, gg = lane >> 4;
        const int phi = 32 * ks + 16 * (j >> 2) + 4 * gg + (j & 3);
        float wf, wb;
        if (lyr < 3) {
            wf = p[lyr * 4096 + (16 * blk + r) * 64 + phi];
            wb = p[lyr * 4096 + phi * 64 + 16 * blk + r];
        } else {
            wf = blk == 0 ? p[12288 + r * 64 + phi] : 0.0f;
            wb = phi < 16 ? p[12288 + phi * 64 + 16 * blk + r] : 0.0f;
        }
        ffwd[e] = (uint16_t)bf16_bits(wf);
        fbwd[e] = (uint16_t)bf16_bits(wb);
    }
}

// One optimizer step of the MLP, one thread per parameter: Adam, then the new weight's bf16
// into its one place in each fragment array (the inverse of k_nrc_adam_mlp's packing map;
// the fragment slots no parameter maps to hold zeros from the last full packing).
__global__ __launch_bounds__(256) void k_nrc_adam_pack(float *__restrict__ p, float *__restrict__ mm,
                                                       float *__restrict__ vv,
                                                       const float *__restrict__ g,
                                                       uint16_t *__restrict__ ffwd,
                                                       uint16_t *__restrict__ fbwd, float lr,
                                                       float b1, float b2, float eps,
                                                       const uint32_t *step, const uint32_t *gmax) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= VG_NRC_MLP_PARAMS) return;
    // a non-finite backward pass (the step's gradient maximum is +inf) changes nothing
    if (!(__uint_as_float(*gmax) <= 3.402823466e38f)) return;
    const float t = (float)*step;
    const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
    float pp = p[i], m = mm[i], v = vv[i];
    adam1(pp, m, v, g[i], lr, b1, b2, eps, c1, c2);
    p[i] = pp;
    mm[i] = m;
    vv[i] = v;
    const uint16_t hb = (uint16_t)bf16_bits(pp);
    // fragment element e = f * 512 + lane * 8 + j with f = lyr * 8 + blk * 2 + ks,
    // lane = gg * 16 + r, phi = 32 ks + 16 (j >> 2) + 4 gg + (j & 3) (k_nrc_adam_mlp)
    auto slot = [](int lyr, int blk, int r, int phi) {
        const int ks = phi >> 5, gg = (phi >> 2) & 3, j = ((phi >> 4) & 1) * 4 + (phi & 3);
        return ((lyr * 8 + blk * 2 + ks) * 64 + gg * 16 + r) * 8 + j;
    };
    if (i < 12288) {
        const int lyr = i >> 12, row = (i >> 6) & 63, col = i & 63;
        ffwd[slot(lyr, row >> 4, row & 15, col)] = hb; // wf = p[lyr][16 blk + r][phi]
        fbwd[slot(lyr, col >> 4, col & 15, row)] = hb; // wb = p[lyr][phi][16 blk + r]
    } else {
        const int i3 = i - 12288, row = i3 >> 6, col = i3 & 63; // row < 16
        ffwd[slot(3, 0, row, col)] = hb;                 // blk 0: p[3][r][phi]
        fbwd[slot(3, col >> 4, col & 15, row)] = hb;     // phi < 16: p[3][phi][16 blk + r]
    }
}

// ---------------------------------------------------------------------- host --
static NrcDev dev_view(const vg_nrc *m) {
    NrcDev d;
    d.grid_h = m->grid_h;
    d.frag_fwd = reinterpret_cast<const u32x4_t *>(m->frag_fwd);
    d.frag_bwd = reinterpret_cast<const u32x4_t *>(m->frag_bwd);
    for (int l = 0; l < VG_NRC_MAX_LEVELS; ++l) d.scale[l] = m->cfg.scale[l];
    d.levels = m->cfg.levels;
    d.log2_table = m->cfg.log2_table;
    return d;
}

static int set_dev(const vg_nrc *m) {
    HIP_TRY(hipSetDevice(m->device));
    return VG_OK;
}

bool nrc_capturing(hipStream_t stream) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return stream && hipStreamIsCapturing(stream, &cs) == hipSuccess &&
           cs == hipStreamCaptureStatusActive;
}

int nrc_join(vg_nrc *m, hipStream_t stream) {
    // A capture does not wait for a step enqueued before it began (the caller joins and
    // synchronises first); steps enqueued inside a capture are joined where they are made.
    if (m->pending && stream != m->tstream && !nrc_capturing(stream))
        HIP_TRY(hipStreamWaitEvent(stream, m->ev_trained, 0));
    return VG_OK;
}

int nrc_infer_launch(vg_nrc *m, const vg_nrc_query *q, size_t q_stride, const uint32_t *sel,
                     const uint32_t *n_dev, uint32_t n_max, float4 *out4, float4 *rad,
                     const float4 *pal, hipStream_t stream) {
    if (n_max == 0) return VG_OK;
    int st = nrc_join(m, stream);
    if (st) return st;
    const uint64_t tiles = ((uint64_t)n_max + 16 * kInferSB - 1) / (16 * kInferSB);
    uint64_t blocks = (tiles + 3) / 4;
    const uint64_t cap = (uint64_t)m->ctx->cus * 8;
    if (blocks > cap) blocks = cap;
    if (rad)
        hipLaunchKernelGGL((k_nrc_infer<true, kInferSB>), dim3((uint32_t)blocks), dim3(256), 0, stream,
                           dev_view(m), q, q_stride, sel, n_dev, n_max, out4, rad, pal);
    else
        hipLaunchKernelGGL((k_nrc_infer<false, kInferSB>), dim3((uint32_t)blocks), dim3(256), 0, stream,
                           dev_view(m), q, q_stride, sel, n_dev, n_max, out4, rad, pal);
    HIP_TRY(hipGetLastError());
    return VG_OK;
}

int nrc_train_launch(vg_nrc *m, const vg_nrc_query *q, const float4 *t, uint32_t n, bool update,
                     hipStream_t stream, float2 *dense) {
    if (n == 0) return VG_OK;
    int jst = nrc_join(m, stream);
    if (jst) return jst;
    if (n > (uint32_t)m->fcap)
        return fail(VG_E_INVALID, "batch of %u above the training buffers' %d", n, m->fcap);
    if (n > (uint32_t)(1u << 31) / (8u * (uint32_t)m->cfg.levels))
        return fail(VG_E_INVALID, "batch of %u: too many grid-gradient records", n);
    const NrcDev dv = dev_view(m);
    const uint32_t eblocks = (uint32_t)(((uint64_t)n * 16u + 255u) / 256u);
    hipLaunchKernelGGL(k_nrc_encode, dim3(eblocks), dim3(256), 0, stream, dv, q, n, m->feat,
                       m->gmax);
    HIP_TRY(hipGetLastError());
    const uint64_t tiles = ((uint64_t)n + kNrcTile - 1) / kNrcTile;
    const int blocks = (int)(tiles < (uint64_t)m->partial_blocks ? tiles : m->partial_blocks);
    hipLaunchKernelGGL(k_nrc_train, dim3(blocks), dim3(256), 0, stream, dv, q, m->feat, t, n,
                       1.0f / (float)n, m->partial, m->dfeat, m->gmax);
    HIP_TRY(hipGetLastError());
    // grid gradient: bucket counts -> scan -> records in bucket order -> per-bucket sums
    const uint32_t nb = (uint32_t)m->gbuckets;
    const size_t nh = (size_t)nb * kGBlocks + 1;
    uint32_t *hoff = m->hist + nh;
    hipLaunchKernelGGL(k_nrc_gcount, dim3(kGBlocks), dim3(kGThreads), nb * sizeof(uint32_t), stream, dv, q,
                       reinterpret_cast<const float2 *>(m->dfeat), n, m->gbits, nb, m->hist);
    HIP_TRY(hipGetLastError());
    HIP_TRY(scan_u32(m->hist, hoff, (uint32_t)nh, static_cast<uint32_t *>(m->ctemp), false,
                     stream));
    hipLaunchKernelGGL(k_nrc_gscatter, dim3(kGBlocks), dim3(kGThreads), nb * sizeof(uint32_t), stream, dv,
                       q, reinterpret_cast<const float2 *>(m->dfeat), n, m->gbits, nb, hoff, m->rk,
                       m->rv);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_nrc_reduce, dim3((VG_NRC_MLP_PARAMS + 1 + 63) / 64), dim3(1024), 0,
                       stream, m->partial, blocks, m->mlp_g, m->loss, m->step, update ? 1 : 0);
    HIP_TRY(hipGetLastError());
    const vg_nrc_config &c = m->cfg;
    hipLaunchKernelGGL(k_nrc_greduce, dim3(nb), dim3(kGRThreads), (2u << m->gbits) * sizeof(uint64_t),
                       stream, hoff, m->rk, m->rv, m->gbits, m->gmax,
                       reinterpret_cast<float2 *>(m->grid), reinterpret_cast<float2 *>(m->grid_m),
                       reinterpret_cast<float2 *>(m->grid_v), m->grid_h, dense, update ? 1 : 0,
                       c.lr_grid, c.beta1, c.beta2, c.eps, m->step);
    HIP_TRY(hipGetLastError());
    if (update) {
        hipLaunchKernelGGL(k_nrc_adam_pack, dim3((VG_NRC_MLP_PARAMS + 255) / 256), dim3(256), 0,
                           stream, m->mlp, m->mlp_m, m->mlp_v, m->mlp_g,
                           reinterpret_cast<uint16_t *>(m->frag_fwd),
                           reinterpret_cast<uint16_t *>(m->frag_bwd), c.lr_mlp, c.beta1, c.beta2,
                           c.eps, m->step, m->gmax);
        HIP_TRY(hipGetLastError());
    }
    return VG_OK;
}

static uint32_t hmix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Uniform in [-1, 1) from (seed, stream, index).
static float unit_sym(uint32_t seed, uint32_t stream, uint64_t i) {
    uint32_t x = hmix32(seed ^ hmix32(stream * 0x9E3779B9u));
    x = hmix32(x ^ (uint32_t)i);
    x = hmix32(x ^ (uint32_t)(i >> 32));
    return (float)((double)(x >> 8) * (2.0 / 16777216.0) - 1.0);
}

static int check_config(const vg_nrc_config *c) {
    if (!c) return fail(VG_E_INVALID, "null config");
    if (c->levels < 1 || c->levels > VG_NRC_MAX_LEVELS)
        return fail(VG_E_INVALID, "levels %d (1..%d)", c->levels, VG_NRC_MAX_LEVELS);
    if (c->log2_table < 4 || c->log2_table > 22)
        return fail(VG_E_INVALID, "log2_table %d (4..22)", c->log2_table);
    for (int l = 0; l < c->levels; ++l)
        if (!(c->scale[l] > 0.0f) || !std::isfinite(c->scale[l]))
            return fail(VG_E_INVALID, "scale[%d] = %g", l, (double)c->scale[l]);
    if (c->batch < 1 || c->batch > VG_NRC_MAX_BATCH)
        return fail(VG_E_INVALID, "batch %d (1..%d)", c->batch, VG_NRC_MAX_BATCH);
    if (c->steps < 0 || c->steps > 64) return fail(VG_E_INVALID, "steps %d (0..64)", c->steps);
    if (!(c->beta1 >= 0.0f && c->beta1 < 1.0f && c->beta2 >= 0.0f && c->beta2 < 1.0f &&
          c->eps >= 0.0f && c->lr_grid >= 0.0f && c->lr_mlp >= 0.0f && c->init_grid >= 0.0f))
        return fail(VG_E_INVALID, "optimizer settings out of range");
    return VG_OK;
}

static void free_grad(vg_nrc *m) {
    void *ptrs[] = {m->rk, m->rv, m->ctemp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    m->rk = nullptr;
    m->rv = nullptr;
    m->ctemp = nullptr;
    m->ctemp_bytes = 0;
}

// Training buffers for batches of up to `cap` samples: encoding, its gradient, and the
// grid-gradient records.
static int alloc_train(vg_nrc *m, int32_t cap) {
    (void)hipFree(m->feat);
    (void)hipFree(m->dfeat);
    m->feat = nullptr;
    m->dfeat = nullptr;
    free_grad(m);
    m->fcap = 0;
    const size_t nc = (size_t)cap * (size_t)m->cfg.levels * 8u;
    bool ok = hipMalloc((void **)&m->feat, (size_t)cap * 16 * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void **)&m->dfeat, (size_t)cap * 32 * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->rk, nc * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void **)&m->rv, nc * sizeof(float2)) == hipSuccess;
    if (!ok) return fail(VG_E_NOMEM, "training buffers for %d samples", cap);
    const size_t nh = (size_t)m->gbuckets * kGBlocks + 1;
    const size_t tb = (size_t)scan_scratch_words((uint32_t)nh) * sizeof(uint32_t);
    if (hipMalloc(&m->ctemp, tb) != hipSuccess)
        return fail(VG_E_NOMEM, "scan scratch of %zu bytes", tb);
    m->ctemp_bytes = tb;
    m->fcap = cap;
    return VG_OK;
}

static void free_nrc(vg_nrc *m) {
    void *ptrs[] = {m->mlp,    m->mlp_m,  m->mlp_v,  m->mlp_g,  m->partial, m->frag_fwd,
                    m->frag_bwd, m->grid, m->grid_m, m->grid_v, m->grid_h,  m->step,
                    m->loss,   m->tq_base, m->tt_base, m->feat, m->dfeat,  m->hist,
                    m->gmax,   m->gdense};
    free_grad(m);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->ev_ready) (void)hipEventDestroy(m->ev_ready);
    if (m->ev_trained) (void)hipEventDestroy(m->ev_trained);
    if (m->tstream) (void)hipStreamDestroy(m->tstream);
}

// Upload parameters, clear the optimizer, refresh the bf16 copies. Synchronous.
static int load_params(vg_nrc *m, const float *mlp, const float *grid) {
    const size_t mb = VG_NRC_MLP_PARAMS * sizeof(float), gb = (size_t)m->n_grid * sizeof(float);
    if (mlp) HIP_TRY(hipMemcpy(m->mlp, mlp, mb, hipMemcpyHostToDevice));
    if (grid) HIP_TRY(hipMemcpy(m->grid, grid, gb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m->mlp_m, 0, mb));
    HIP_TRY(hipMemset(m->mlp_v, 0, mb));
    HIP_TRY(hipMemset(m->mlp_g, 0, mb));
    HIP_TRY(hipMemset(m->grid_m, 0, gb));
    HIP_TRY(hipMemset(m->grid_v, 0, gb));
    HIP_TRY(hipMemset(m->step, 0, sizeof(uint32_t)));
    HIP_TRY(hipMemset(m->loss, 0, sizeof(float)));
    hipLaunchKernelGGL(k_nrc_grid_pack, dim3((uint32_t)((m->n_entries + 255) / 256)), dim3(256), 0,
                       nullptr, reinterpret_cast<const float2 *>(m->grid), m->grid_h, m->n_entries);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_nrc_adam_mlp, dim3(1), dim3(1024), 0, nullptr, m->mlp, m->mlp_m, m->mlp_v,
                       m->mlp_g, reinterpret_cast<uint16_t *>(m->frag_fwd),
                       reinterpret_cast<uint16_t *>(m->frag_bwd), 0.0f, 0.0f, 0.0f, 0.0f, m->step, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return VG_OK;
}

// Validates a host-called batch and grows the training buffers to hold it (synchronous
// reallocation: a batch above the configured size must not be enqueued during capture).
static int check_batch(vg_nrc *m, const vg_nrc_query *q, const float *t, uint32_t n) {
    if (!m) return fail(VG_E_INVALID, "null cache");
    if (n > VG_NRC_MAX_BATCH) return fail(VG_E_INVALID, "n %u above %d", n, VG_NRC_MAX_BATCH);
    if (n && (!q || !t)) return fail(VG_E_INVALID, "null queries or targets");
    int st = set_dev(m);
    if (st) return st;
    if (n > (uint32_t)m->fcap) {
        HIP_TRY(hipDeviceSynchronize());
        st = alloc_train(m, (int32_t)n);
        if (st) return st;
    }
    return VG_OK;
}

} // namespace vg

using namespace vg;

extern "C" {

int vg_nrc_scales(int32_t levels, float coarse_cell, float fine_cell, float *out) {
    if (!out || levels < 1 || levels > VG_NRC_MAX_LEVELS || !(coarse_cell > 0.0f) ||
        !(fine_cell > 0.0f))
        return fail(VG_E_INVALID, "bad scale arguments");
    const double b = levels > 1 ? std::pow((double)coarse_cell / fine_cell, 1.0 / (levels - 1)) : 1.0;
    for (int l = 0; l < levels; ++l) out[l] = (float)(std::pow(b, l) / (double)coarse_cell);
    return VG_OK;
}

int vg_nrc_config_default(vg_nrc_config *c) {
    if (!c) return fail(VG_E_INVALID, "null config");
    std::memset(c, 0, sizeof(*c));
    c->levels = 16;
    c->log2_table = 19;
    vg_nrc_scales(16, 128.0f, 1.0f, c->scale);
    c->lr_grid = 1e-2f;
    c->lr_mlp = 1e-2f;
    c->beta1 = 0.9f;
    c->beta2 = 0.99f;
    c->eps = 1e-10f;
    c->init_grid = 1e-4f;
    c->seed = 0x5EEDCAC4u;
    c->batch = 65536;
    c->steps = 1;
    return VG_OK;
}

int vg_nrc_create(vg_ctx *ctx, const vg_nrc_config *cfg, vg_nrc **out) {
    if (!ctx || !out) return fail(VG_E_INVALID, "null argument");
    *out = nullptr;
    int st = check_config(cfg);
    if (st) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    vg_nrc *m = new vg_nrc();
    std::memset(m, 0, sizeof(*m));
    m->ctx = ctx;
    m->device = ctx->device;
    m->cfg = *cfg;
    m->n_entries = (int64_t)cfg->levels << cfg->log2_table;
    m->n_grid = 2 * m->n_entries;
    m->partial_blocks = 1024;
    m->tcap = cfg->batch;
    m->gbits = cfg->log2_table < 12 ? cfg->log2_table : 12;
    m->gbuckets = cfg->levels << (cfg->log2_table - m->gbits);
    const size_t mb = VG_NRC_MLP_PARAMS * sizeof(float), gb = (size_t)m->n_grid * sizeof(float);
    const size_t fb = (size_t)kNrcFrags * 64 * 16;
    bool ok = hipMalloc((void **)&m->mlp, mb) == hipSuccess &&
              hipMalloc((void **)&m->mlp_m, mb) == hipSuccess &&
              hipMalloc((void **)&m->mlp_v, mb) == hipSuccess &&
              hipMalloc((void **)&m->mlp_g, mb) == hipSuccess &&
              hipMalloc((void **)&m->partial,
                        (size_t)m->partial_blocks * kNrcPartial * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&m->frag_fwd, fb) == hipSuccess &&
              hipMalloc((void **)&m->frag_bwd, fb) == hipSuccess &&
              hipMalloc((void **)&m->grid, gb) == hipSuccess &&
              hipMalloc((void **)&m->grid_m, gb) == hipSuccess &&
              hipMalloc((void **)&m->grid_v, gb) == hipSuccess &&
              hipMalloc((void **)&m->grid_h, (size_t)m->n_entries * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void **)&m->hist,
                        2 * ((size_t)m->gbuckets * kGBlocks + 1) * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc((void **)&m->gmax, 256) == hipSuccess &&
              hipMalloc((void **)&m->step, 256) == hipSuccess &&
              hipMalloc((void **)&m->loss, 256) == hipSuccess &&
              hipMalloc((void **)&m->tq_base, 2 * (size_t)m->tcap * sizeof(vg_nrc_query)) == hipSuccess &&
              hipMalloc((void **)&m->tt_base, 2 * (size_t)m->tcap * sizeof(float4)) == hipSuccess &&
              hipStreamCreateWithFlags(&m->tstream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&m->ev_ready, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&m->ev_trained, hipEventDisableTiming) == hipSuccess;
    m->tq = m->tq_base;
    m->tt = m->tt_base;
    if (ok) ok = alloc_train(m, cfg->batch > 65536 ? cfg->batch : 65536) == VG_OK;
    if (!ok) {
        free_nrc(m);
        delete m;
        return fail(VG_E_NOMEM, "radiance cache storage (%lld grid parameters)",
                    (long long)(2 * ((int64_t)cfg->levels << cfg->log2_table)));
    }
    // Xavier-uniform MLP (W3 rows >= 3 zero), uniform grid features.
    std::vector<float> mlp(VG_NRC_MLP_PARAMS, 0.0f), grid((size_t)m->n_grid);
    const float s_hidden = std::sqrt(6.0f / 128.0f), s_out = std::sqrt(6.0f / 67.0f);
    for (int i = 0; i < 3 * 4096; ++i) mlp[i] = unit_sym(cfg->seed, 1, (uint64_t)i) * s_hidden;
    for (int i = 0; i < 3 * 64; ++i) mlp[12288 + i] = unit_sym(cfg->seed, 2, (uint64_t)i) * s_out;
    for (int64_t i = 0; i < m->n_grid; ++i) grid[(size_t)i] = unit_sym(cfg->seed, 3, (uint64_t)i) * cfg->init_grid;
    HIP_TRY(hipMemset(m->tq_base, 0, 2 * (size_t)m->tcap * sizeof(vg_nrc_query)));
    HIP_TRY(hipMemset(m->tt_base, 0, 2 * (size_t)m->tcap * sizeof(float4)));
    st = load_params(m, mlp.data(), grid.data());
    if (st) {
        free_nrc(m);
        delete m;
        return st;
    }
    *out = m;
    return VG_OK;
}

int vg_nrc_destroy(vg_nrc *m) {
    if (!m) return VG_OK;
    (void)hipSetDevice(m->device);
    (void)hipDeviceSynchronize();
    free_nrc(m);
    delete m;
    return VG_OK;
}

int vg_nrc_param_counts(const vg_nrc *m, int64_t out[2]) {
    if (!m || !out) return fail(VG_E_INVALID, "null argument");
    out[0] = VG_NRC_MLP_PARAMS;
    out[1] = m->n_grid;
    return VG_OK;
}

int vg_nrc_get_params(vg_nrc *m, float *mlp, float *grid) {
    if (!m) return fail(VG_E_INVALID, "null cache");
    int st = set_dev(m);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    if (mlp) HIP_TRY(hipMemcpy(mlp, m->mlp, VG_NRC_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
    if (grid)
        HIP_TRY(hipMemcpy(grid, m->grid, (size_t)m->n_grid * sizeof(float), hipMemcpyDeviceToHost));
    return VG_OK;
}

int vg_nrc_set_params(vg_nrc *m, const float *mlp, const float *grid) {
    if (!m) return fail(VG_E_INVALID, "null cache");
    int st = set_dev(m);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    return load_params(m, mlp, grid);
}

int vg_nrc_infer(vg_nrc *m, const vg_nrc_query *q, uint32_t n, float *out4, void *stream) {
    if (!m) return fail(VG_E_INVALID, "null cache");
    if (n && (!q || !out4)) return fail(VG_E_INVALID, "null queries or output");
    int st = set_dev(m);
    if (st) return st;
    return nrc_infer_launch(m, q, 0, nullptr, nullptr, n, reinterpret_cast<float4 *>(out4),
                            nullptr, nullptr, (hipStream_t)stream);
}

int vg_nrc_join(vg_nrc *m, void *stream) {
    if (!m) return fail(VG_E_INVALID, "null cache");
    int st = set_dev(m);
    if (st) return st;
    return nrc_join(m, (hipStream_t)stream);
}

int vg_nrc_train_step(vg_nrc *m, const vg_nrc_query *q, const float *t, uint32_t n, void *stream) {
    int st = check_batch(m, q, t, n);
    if (st) return st;
    if (n == 0) return fail(VG_E_INVALID, "empty batch");
    return nrc_train_launch(m, q, reinterpret_cast<const float4 *>(t), n, true, (hipStream_t)stream);
}

int vg_nrc_gradients(vg_nrc *m, const vg_nrc_query *q, const float *t, uint32_t n,
                     float *mlp_grad, float *grid_grad, float *loss) {
    int st = check_batch(m, q, t, n);
    if (st) return st;
    if (n == 0) return fail(VG_E_INVALID, "empty batch");
    HIP_TRY(hipDeviceSynchronize());
    if (grid_grad && !m->gdense &&
        hipMalloc((void **)&m->gdense, (size_t)m->n_entries * sizeof(float2)) != hipSuccess)
        return fail(VG_E_NOMEM, "dense grid gradient");
    st = nrc_train_launch(m, q, reinterpret_cast<const float4 *>(t), n, false, nullptr,
                          grid_grad ? m->gdense : nullptr);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    if (mlp_grad)
        HIP_TRY(hipMemcpy(mlp_grad, m->mlp_g, VG_NRC_MLP_PARAMS * sizeof(float),
                          hipMemcpyDeviceToHost));
    if (grid_grad)
        HIP_TRY(hipMemcpy(grid_grad, m->gdense, (size_t)m->n_grid * sizeof(float),
                          hipMemcpyDeviceToHost));
    if (loss) HIP_TRY(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipDeviceSynchronize());
    return VG_OK;
}

int vg_nrc_stats(vg_nrc *m, int64_t *steps, float *loss) {
    if (!m) return fail(VG_E_INVALID, "null cache");
    int st = set_dev(m);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    uint32_t s = 0;
    float l = 0.0f;
    HIP_TRY(hipMemcpy(&s, m->step, sizeof(s), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&l, m->loss, sizeof(l), hipMemcpyDeviceToHost));
    if (steps) *steps = s;
    if (loss) *loss = l;
    return VG_OK;
}

int vg_nrc_last_batch(vg_nrc *m, vg_nrc_query *q, float *t) {
    if (!m) return fail(VG_E_INVALID, "null cache");
    int st = set_dev(m);
    if (st) return st;
    HIP_TRY(hipDeviceSynchronize());
    if (q) HIP_TRY(hipMemcpy(q, m->tq, (size_t)m->tcap * sizeof(vg_nrc_query), hipMemcpyDeviceToHost));
    if (t) HIP_TRY(hipMemcpy(t, m->tt, (size_t)m->tcap * sizeof(float4), hipMemcpyDeviceToHost));
    return VG_OK;
}

} // extern "C"
