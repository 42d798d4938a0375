// Internal declarations shared by the host builders (vg_host.cpp) and the device
// engine (vg_device.hip). Not part of the ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/vg.h"

namespace vg {

// Set the thread-local error message and return `status`.
int fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Geometry of a window: counts per level and axis (src/world_dims.rs:9-37).
struct Dims {
    int32_t v[3];  // voxels
    int32_t b[3];  // bricks (4^3 voxels)
    int32_t t[3];  // tiles (16^3 voxels)
    int32_t c[3];  // chunks (64^3 voxels)
    int32_t l4[3]; // L4 cells (256^3 voxels)
    int64_t n_bricks, n_tiles, n_chunks, n_l4;
};

// Validates dims (positive multiples of 64, each below 2^15 so every plane
// coordinate is exact in fp32) and fills `out`.
int make_dims(const int32_t dims[3], Dims *out);

inline int64_t lin3(const int32_t n[3], int32_t x, int32_t y, int32_t z) {
    return (int64_t)x + (int64_t)y * n[0] + (int64_t)z * n[0] * n[1];
}
// Child bit of a 4^3 sub-cell inside its parent u64, every level: lx + 4 lz + 16 ly.
inline int child_bit(int32_t lx, int32_t ly, int32_t lz) { return lx + lz * 4 + ly * 16; }
inline int32_t pos_mod(int32_t a, int32_t b) {
    int32_t r = a % b;
    return r < 0 ? r + b : r;
}

int resolve_threads(int threads);

} // namespace vg
