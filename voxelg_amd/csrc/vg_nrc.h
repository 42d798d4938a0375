// Radiance-cache state shared by the cache kernels (vg_nrc.hip) and the learned probe
// update (vg_gi.hip). Not part of the ABI.
#pragma once

#include "vg_world.h"

struct vg_nrc {
    vg_ctx *ctx;
    int device;             // ctx's device (destroy does not touch ctx)
    vg_nrc_config cfg;
    int64_t n_entries;      // levels * 2^log2_table
    int64_t n_grid;         // grid parameters = 2 * n_entries
    float *mlp, *mlp_m, *mlp_v, *mlp_g;  // fp32 master, Adam moments, reduced gradient
    float *partial;         // per block: VG_NRC_MLP_PARAMS gradient sums + loss
    int partial_blocks;     // blocks the partial buffer holds
    uint32_t *frag_fwd;     // bf16 MFMA A-operand fragments of W (4 layers x 4 x 2 x 64 lanes x 16 B)
    uint32_t *frag_bwd;     // ... of W^T for the backward pass
    float *grid, *grid_m, *grid_v;  // fp32 master and Adam moments
    uint32_t *grid_h;       // bf16 feature pairs the kernels read, one u32 per entry
    uint32_t *step;         // optimizer steps taken (device)
    float *loss;            // loss of the last step (device)
    vg_nrc_query *tq;       // the learned probe update's last training batch
    float4 *tt;             // ... and its targets
    int32_t tcap;           // entries of tq / tt
    // Training batches are double-buffered: a learned update's path launch writes one half
    // while the previous update's training step, still running on tstream, reads the other.
    vg_nrc_query *tq_base;  // 2 * tcap queries (halves at 0 and tcap)
    float4 *tt_base;        // 2 * tcap targets
    int tpar;               // half written by the last learned update
    hipStream_t tstream;    // the learned update's training step overlaps the next update
    hipEvent_t ev_ready;    // the update's cache inference has run (parameters may change)
    hipEvent_t ev_trained;  // the last training step enqueued on tstream has run
    bool pending;           // ev_trained has been recorded
    uint32_t *feat;         // training batch encoding: bf16 pair per (sample, level)
    float *dfeat;           // ... its gradient: 32 fp32 per sample
    int32_t fcap;           // samples the training buffers hold
    // grid gradient (samples x levels x 8 corner records): records grouped by entry bucket
    // (2^gbits entries each) through per-block bucket counts and their exclusive scan, then
    // summed per bucket in LDS in 64-bit fixed point (exact, so any order gives the same sums)
    uint32_t *rk;           // record: entry index inside its bucket
    float2 *rv;             // record: w * d loss / d feature pair
    uint32_t *hist;         // per (bucket, count block) record counts, scanned in place
    uint32_t *gmax;         // max |d loss / d feature| of the step (fp32 bits)
    float2 *gdense;         // dense grid gradient, for vg_nrc_gradients only (lazy)
    void *ctemp;            // scan scratch
    size_t ctemp_bytes;
    int gbits;              // log2 entries per bucket
    int gbuckets;           // levels << (log2_table - gbits)
};

namespace vg {

constexpr int kNrcWidth = 64;
constexpr int kNrcFrags = 32;       // [layer 4][block 4][k-step 2]
constexpr int kNrcTile = 256;       // training samples per block per tile
constexpr int kNrcPartial = VG_NRC_MLP_PARAMS + 16; // floats per block partial (loss at +0)

// Inference over the queries at q + sel[0] * q_stride (sel may be null: q itself), n_dev[0]
// of them (else n_max; never more than n_max). out4 receives raw predictions; with rad/pal
// set (probe epilogue) the kernel instead adds albedo * max(0, prediction) to rad[aux]
// (aux = sample id, face >> 8 = material).
int nrc_infer_launch(vg_nrc *m, const vg_nrc_query *q, size_t q_stride, const uint32_t *sel,
                     const uint32_t *n_dev, uint32_t n_max, float4 *out4, float4 *rad,
                     const float4 *pal, hipStream_t stream);
// Loss gradient of n queries: the MLP gradient reduced to mlp_g, the grid gradient summed
// per entry; with update, the optimizer step over both follows; with dense, the grid
// gradient of every entry is also written there (n_entries float pairs).
int nrc_train_launch(vg_nrc *m, const vg_nrc_query *q, const float4 *t, uint32_t n, bool update,
                     hipStream_t stream, float2 *dense = nullptr);
// Make `stream` wait for the training step still running on the cache's own stream (work
// that reads or writes the cache parameters must come after it). Under graph capture it
// does not wait for a step enqueued before the capture began.
int nrc_join(vg_nrc *m, hipStream_t stream);
// `stream` is being captured into a HIP graph.
bool nrc_capturing(hipStream_t stream);

} // namespace vg
