"""Pin the oracle's noise, fBm and seed offsets (src/voxel.rs:1163-1232, 1999-2021)."""
import numpy as np

import oracle

M32 = 0xFFFFFFFF


def py_hash2(x, z):
    # the same u32 wrapping arithmetic in arbitrary-precision integers
    h = ((x & M32) * 0x9E3779B1 + (z & M32) * 0x85EBCA77) & M32
    h = (h * 0xC2B2AE3D) & M32
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    return np.float32(h & 0xFFFFFF) / np.float32(0xFFFFFF) * np.float32(2) - np.float32(1)


def py_hash3(x, y, z):
    h = ((x & M32) * 0x9E3779B1 + (y & M32) * 0x85EBCA77 + (z & M32) * 0xC2B2AE3D) & M32
    h = (h * 0xD2B74407) & M32
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    return np.float32(h & 0xFFFFFF) / np.float32(0xFFFFFF) * np.float32(2) - np.float32(1)


def test_hashes_wrap_like_u32():
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(-2**31, 2**31 - 1, (200, 3)):
        assert oracle.hash2(x, z) == py_hash2(int(x), int(z))
        assert oracle.hash3(x, y, z) == py_hash3(int(x), int(y), int(z))
        assert -1.0 <= oracle.hash2(x, z) <= 1.0


def test_value_noise_lattice_and_midpoints():
    rng = np.random.default_rng(1)
    for x, z in rng.integers(-1000, 1000, (100, 2)):
        # smoothstep(0) = 0: a lattice point returns its own hash exactly
        assert oracle.value_noise_2d(float(x), float(z)) == oracle.hash2(x, z)
        assert oracle.value_noise_3d(float(x), float(z), 3.0) == oracle.hash3(x, z, 3)
        # smoothstep(1/2) = 1/2: the midpoint of an edge is the mean of its ends
        mid = oracle.value_noise_2d(x + 0.5, float(z))
        assert abs(mid - 0.5 * (oracle.hash2(x, z) + oracle.hash2(x + 1, z))) < 1e-6


def test_value_noise_continuous_across_cells():
    for x in np.linspace(-3, 3, 61):
        a = oracle.value_noise_2d(np.float32(x) - np.float32(1e-4), 0.37)
        b = oracle.value_noise_2d(np.float32(x) + np.float32(1e-4), 0.37)
        assert abs(a - b) < 1e-2
        a = oracle.value_noise_3d(0.2, np.float32(x) - np.float32(1e-4), 0.9)
        b = oracle.value_noise_3d(0.2, np.float32(x) + np.float32(1e-4), 0.9)
        assert abs(a - b) < 1e-2


def test_fbm():
    for x, z in [(0.3, 0.7), (10.25, -4.5), (-100.1, 3.3)]:
        assert oracle.fbm_2d(x, z, 1) == oracle.value_noise_2d(x, z)
        v2 = (oracle.value_noise_2d(x, z) + 0.5 * oracle.value_noise_2d(2 * x, 2 * z)) / 1.5
        assert abs(oracle.fbm_2d(x, z, 2) - v2) < 1e-6
        assert -1.0 <= oracle.fbm_2d(x, z, 5) <= 1.0


def test_seed_offsets():
    hi, lo = 0xFFFFFFFE, 7  # hi reads as i32 -2
    s = oracle.seed_offset_xz((hi << 32) | lo)
    assert s == (float(np.float32(-2.0) * np.float32(0.01734)),
                 float(np.float32(7.0) * np.float32(0.02153)))
