// Prefix sums of u32 counts on the device, for the passes that place records by bucket (the
// radiance cache's grid gradient, the sphere edits' brick runs). Three launches: each tile of
// 4096 values is scanned by one block, one block scans the tile totals, and a last pass adds
// each tile's offset. Fixed summation order (deterministic), no host synchronisation and no
// memset, so a HIP graph may capture it. Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vg {
namespace { // each source file that includes this header gets its own kernels

constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanPer = 4;                       // values per thread
constexpr uint32_t kScanTile = kScanThreads * kScanPer; // values per block
constexpr uint32_t kScanMaxTiles = kScanTile;          // tile totals one block can scan

// Exclusive scan of v over the block's threads (kScanThreads); returns the block total in
// `total`. Uses `ws` (kScanThreads / 64 words of LDS).
__device__ __forceinline__ uint32_t scan_block_exclusive(uint32_t v, uint32_t *ws, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v; // inclusive scan within the wave
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63u) ws[wave] = x;
    __syncthreads();
    if (wave == 0) {
        const uint32_t nw = kScanThreads / 64u;
        uint32_t s = lane < nw ? ws[lane] : 0u;
#pragma unroll
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t y = __shfl_up(s, off, 64);
            if (lane >= off) s += y;
        }
        if (lane < nw) ws[lane] = s; // inclusive wave totals
    }
    __syncthreads();
    total = ws[kScanThreads / 64u - 1u];
    const uint32_t before = wave ? ws[wave - 1u] : 0u;
    __syncthreads(); // ws is reused by the caller's next scan
    return before + x - v;
}

// Tile t of `in` (n values) scanned into `out` (exclusive, or inclusive with `incl`), its
// total into sums[t].
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint32_t *in, uint32_t *out,
                                                             uint32_t n, uint32_t *sums, int incl) {
    __shared__ uint32_t ws[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
        v[j] = base + j < n ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = scan_block_exclusive(s, ws, total);
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
        if (base + j < n) out[base + j] = incl ? run + v[j] : run;
        run += v[j];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the tile totals in place (one block, at most kScanMaxTiles).
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(uint32_t *sums, uint32_t ntiles) {
    __shared__ uint32_t ws[kScanThreads / 64];
    const uint32_t base = threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
        v[j] = base + j < ntiles ? sums[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t run = scan_block_exclusive(s, ws, total);
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j) {
        if (base + j < ntiles) sums[base + j] = run;
        run += v[j];
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t *__restrict__ out, uint32_t n,
                                                           const uint32_t *__restrict__ sums) {
    const uint32_t add = sums[blockIdx.x];
    if (blockIdx.x == 0 || add == 0u) return;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; ++j)
        if (base + j < n) out[base + j] += add;
}

// Scratch words scan_u32 needs for n values.
inline uint32_t scan_scratch_words(uint32_t n) { return (n + kScanTile - 1) / kScanTile; }

// out = exclusive (incl = false) or inclusive prefix sums of in[0, n); in == out is allowed.
// `sums` holds scan_scratch_words(n) words. n <= kScanTile * kScanMaxTiles (2^24).
inline hipError_t scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *sums,
                           bool incl, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t tiles = scan_scratch_words(n);
    if (tiles > kScanMaxTiles) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, stream, in, out, n, sums,
                       incl ? 1 : 0);
    if (tiles > 1) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, stream, sums, tiles);
        hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanThreads), 0, stream, out, n, sums);
    }
    return hipGetLastError();
}

} // namespace
} // namespace vg
