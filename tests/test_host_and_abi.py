"""CPU-only checks of the product library's host side and its C ABI (no GPU calls)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import oracle
import voxelg_amd as vg
from voxelg_amd import workloads
from voxelg_amd._lib import IT_BYTES, LIB_PATH, VGError, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(vg_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 30
    L = C.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert lib().vg_abi_version() == 1


def test_status_strings_and_errors():
    L = lib()
    assert L.vg_status_string(0) == b"ok"
    with pytest.raises(VGError) as e:
        vg.world_counts((100, 64, 64))
    assert "multiple of 64" in str(e.value)
    assert e.value.status == -1


def test_counts_match_reference_dims():
    # src/world_dims.rs:9-37 at the reference window 512x256x512
    nb, nt, nc, nl4, nv = vg.world_counts((512, 256, 512))
    assert (nb, nt, nc, nl4, nv) == (128 * 64 * 128, 32 * 16 * 32, 8 * 4 * 8, 2 * 1 * 2,
                                     512 * 256 * 512)
    assert vg.world_counts((2048, 256, 2048))[:4] == (16777216, 262144, 4096, 64)
    assert oracle.counts((2048, 256, 2048))[:4] == (16777216, 262144, 4096, 64)


def test_generator_matches_oracle_bit_exact():
    seed = workloads.SMALL.seed
    w = vg.World((256, 128, 256)).generate(seed)
    ow = oracle.OracleWorld((256, 128, 256)).generate(seed)
    assert w.bricks.tobytes() == ow.bricks.tobytes()
    for a in ("tile_mask", "chunk_mask", "l4_mask", "brick_uniform", "tile_uniform"):
        assert np.array_equal(getattr(w, a), getattr(ow, a)), a
    # shifted window: same world voxels, folded into toroidal slots
    o = (256, 0, -256)
    w2 = vg.World((256, 128, 256), origin=o).generate(seed)
    ow2 = oracle.OracleWorld((256, 128, 256), origin=o).generate(seed)
    assert w2.bricks.tobytes() == ow2.bricks.tobytes()


def test_generator_invariants():
    seed = 12345
    dims = (256, 128, 256)
    w = vg.World(dims).generate(seed)
    rng = np.random.default_rng(0)
    for x, z in rng.integers(0, 256, (40, 2)):
        h = vg.world.terrain_height(dims, seed, x, z)
        assert h == oracle.terrain_height(dims, seed, x, z)
        assert 2 <= h <= 127
        col = w.material_at(np.array([[x, y, z] for y in range(128)]))
        assert col[h] == (1 if h <= 57 else 2)          # sand / grass on top
        assert all(col[y] in (0, 3) for y in range(max(0, h - 4), h))  # dirt (or cave)
        assert all(col[y] == 12 for y in range(h + 1, 56))  # water below y = 56
        assert all(col[y] == 0 for y in range(max(h + 1, 56), 128))
        assert all(col[y] in (0, 4) for y in range(5, max(5, h - 4)))
        assert all(col[y] == 4 for y in range(0, min(5, h - 4)))  # no caves in the bottom 5
    a = vg.World(dims).generate(seed)
    b = vg.World(dims).generate(seed + 1)
    assert a.bricks.tobytes() == w.bricks.tobytes()  # deterministic
    assert a.bricks.tobytes() != b.bricks.tobytes()  # seed-dependent
    assert 0.2 < w.solid_fraction() < 0.8


def test_pyramid_matches_oracle_and_brute_force():
    rng = np.random.default_rng(1)
    dims = (128, 64, 192)
    w = vg.World(dims)
    xyz = rng.integers(0, [128, 64, 192], (3000, 3))
    w.set_voxels(xyz, rng.integers(1, 13, 3000), rebuild=False)
    # a fully solid uniform tile and brick
    blk = np.stack(np.meshgrid(np.arange(16, 32), np.arange(16, 32), np.arange(32, 48),
                               indexing="ij"), -1).reshape(-1, 3)
    w.set_voxels(blk, 4, rebuild=False)
    w.build_pyramid()
    ow = oracle.OracleWorld(dims)
    ow.bricks[:] = w.bricks
    ow.rebuild()
    for a in ("tile_mask", "chunk_mask", "l4_mask", "brick_uniform", "tile_uniform"):
        assert np.array_equal(getattr(w, a), getattr(ow, a)[: len(getattr(w, a))]), a
    # brute force: bit set <=> child non-empty
    bx, by, bz = 32, 16, 48
    occ = (w.bricks["occupancy"] != 0).reshape(bz, by, bx)
    for t in range(len(w.tile_mask)):
        tx, ty, tz = t % 8, (t // 8) % 4, t // 32
        sub = occ[tz * 4:tz * 4 + 4, ty * 4:ty * 4 + 4, tx * 4:tx * 4 + 4]
        expect = 0
        for lz in range(4):
            for ly in range(4):
                for lx in range(4):
                    if sub[lz, ly, lx]:
                        expect |= 1 << (lx + 4 * lz + 16 * ly)
        assert int(w.tile_mask[t]) == expect
    assert (w.tile_uniform == 4).sum() == 1 and (w.brick_uniform == 4).sum() >= 64


def test_incremental_set_voxel_equals_full_rebuild():
    # src/voxel.rs:879-946 incremental mask/uniform upkeep == rebuild_all_* (oracle)
    rng = np.random.default_rng(2)
    ow = oracle.OracleWorld((128, 64, 128))
    for _ in range(4000):
        x, y, z = rng.integers(0, [128, 64, 128])
        m = int(rng.integers(0, 5)) if rng.random() < 0.7 else 0
        ow.set_voxel(x, y, z, m)
    ref = oracle.OracleWorld((128, 64, 128))
    ref.bricks[:] = ow.bricks
    ref.rebuild()
    for a in ("tile_mask", "chunk_mask", "l4_mask", "brick_uniform", "tile_uniform"):
        assert np.array_equal(getattr(ow, a), getattr(ref, a)), a


def test_brick_uniform_rule():
    b = np.zeros(1, oracle.BRICK_DTYPE)
    b["occupancy"] = np.uint64(0xFFFFFFFFFFFFFFFF)
    b["materials"] = 7
    assert oracle.brick_uniform_of(b) == 7
    b["materials"][0, 5] = 8
    assert oracle.brick_uniform_of(b) == 0
    b["materials"] = 7
    b["occupancy"] = np.uint64(0xFFFFFFFFFFFFFFFE)
    assert oracle.brick_uniform_of(b) == 0


def test_camera_basis_closed_form():
    u = vg.Camera(pos=(1, 2, 3), yaw=0.0, pitch=0.0, fov_y=math.radians(90)).uniform(64, 32)
    assert np.allclose(u[4:7], [0, 0, 1]) and np.allclose(u[8:11], [1, 0, 0])
    assert np.allclose(u[12:15], [0, 1, 0]) and abs(u[15] - 1.0) < 1e-6
    assert list(u[16:18]) == [64.0, 32.0] and list(u[0:3]) == [1.0, 2.0, 3.0]
    rng = np.random.default_rng(3)
    for _ in range(20):
        c = vg.Camera(pos=(0, 0, 0), yaw=rng.uniform(0, 6.28), pitch=rng.uniform(-1.5, 1.5))
        u = c.uniform(10, 10)
        f, r, up = u[4:7], u[8:11], u[12:15]
        for v in (f, r, up):
            assert abs(np.linalg.norm(v) - 1) < 1e-5
        assert abs(f @ r) < 1e-5 and abs(f @ up) < 1e-5 and abs(r @ up) < 1e-5
        assert abs(r[1]) < 1e-7  # right stays horizontal (src/camera.rs:35-39)
    # the centre of an odd image looks straight down forward
    cam = vg.Camera(pos=(0, 0, 0), yaw=0.4, pitch=-0.2).uniform(101, 51)
    d = oracle.primary_dirs(cam, 101, 51)[25 * 101 + 50]
    assert np.allclose(d, cam[4:7], atol=1e-6) and abs(np.linalg.norm(d) - 1) < 1e-6


def test_byte_table_is_shared():
    assert tuple(IT_BYTES) == tuple(oracle.IT_BYTES)


def test_device_entry_points_fail_cleanly_without_gpu():
    from voxelg_amd.device import device_count
    if device_count() > 0:
        pytest.skip("a device is present")
    from voxelg_amd.device import Context
    with pytest.raises(VGError) as e:
        Context(0)
    assert e.value.status == -3
