"""Pin R1, the exact hierarchical walk the GPU reproduces bit for bit:
  - R1 == the same walk one voxel per iteration (no hierarchy, no uniform shortcut)
  - R1 == a brute-force nearest-slab search over every solid voxel on tiny worlds
  - R1 == R0 (src/raycast.rs restated) on every clean hit of a benchmark frame
  - LOD (shaders/raymarch.wgsl:4673-4758) on hand-built bricks with known answers."""
import numpy as np
import pytest

import oracle
import scenes
from voxelg_amd import workloads


def random_rays(rng, n, dims, inside_frac=0.7, tmax=np.inf):
    dims = np.asarray(dims, np.float32)
    o = rng.uniform(0, 1, (n, 3)).astype(np.float32) * dims
    out = rng.random(n) > inside_frac
    o[out] = (rng.uniform(-0.5, 1.5, (out.sum(), 3)) * dims).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    # axis-aligned, tiny and exactly-zero components
    k = rng.integers(0, 3, n)
    sel = rng.random(n) < 0.15
    d[sel, k[sel]] = 0.0
    sel = rng.random(n) < 0.05
    d[sel, k[sel]] = np.float32(3e-9)
    sel = rng.random(n) < 0.05
    d[sel] = np.eye(3, dtype=np.float32)[k[sel]] * rng.choice([-1, 1], sel.sum())[:, None]
    return oracle.make_rays(o, d, tmax)


def same(a, b):
    return a.tobytes() == b.tobytes()


def first_diff(a, b):
    i = np.nonzero(a.view(np.uint8).reshape(len(a), -1).any(1) !=
                   b.view(np.uint8).reshape(len(b), -1).any(1))[0]
    m = np.nonzero((a.view(np.uint8).reshape(len(a), -1) != b.view(np.uint8).reshape(len(b), -1)).any(1))[0]
    return m[:5]


@pytest.mark.parametrize("flags", [0, oracle.SKIP_TRANSPARENT, oracle.ANY_HIT])
def test_r1_equals_flat_walk(small_world, flags):
    ow = oracle.OracleWorld.view(small_world)
    rng = np.random.default_rng(7 + flags)
    rays = random_rays(rng, 3000, ow.dims, tmax=np.float32(300.0))
    got, _ = ow.trace_rays(rays, flags)
    ref = ow.trace_flat(rays, flags)
    bad = first_diff(got, ref)
    assert same(got, ref), (bad, got[bad], ref[bad])
    assert got["hit"].sum() > 500


def test_r1_equals_flat_walk_shifted_origin():
    ow = scenes.floor_pillars((96, 0, 160))
    rng = np.random.default_rng(3)
    n = 2000
    base = np.array([96, 0, 160], np.float32)
    o = base + rng.uniform([0, 60, 0], [80, 140, 80], (n, 3)).astype(np.float32)
    tgt = base + rng.uniform([4, 50, 4], [70, 100, 70], (n, 3)).astype(np.float32)
    rays = oracle.make_rays(o, tgt - o)
    got, _ = ow.trace_rays(rays)
    ref = ow.trace_flat(rays)
    assert same(got, ref)
    assert got["hit"].sum() > 500


def tiny_world(rng, density):
    ow = oracle.OracleWorld((64, 64, 64))
    n = int(64 ** 3 * density)
    xyz = rng.integers(0, 64, (n, 3))
    mats = rng.integers(1, 13, n)
    for (x, y, z), m in zip(xyz, mats):
        ow.set_voxel(x, y, z, m)
    return ow


@pytest.mark.parametrize("density", [0.002, 0.02])
def test_r1_equals_brute_force(density):
    rng = np.random.default_rng(int(density * 1e4))
    ow = tiny_world(rng, density)
    rays = random_rays(rng, 1000, ow.dims, inside_frac=0.8)
    # start in air, away from voxel faces, so first-hit is unambiguous
    keep = []
    for i, r in enumerate(rays):
        v = np.floor(r["o"]).astype(int)
        inside = np.all((v >= 0) & (v < 64))
        if inside and ow.material_at(*v):
            continue
        keep.append(i)
    rays = rays[keep]
    got, _ = ow.trace_rays(rays)
    ref, ties = ow.trace_brute(rays)
    n_checked = 0
    for i in range(len(rays)):
        if ties[i] > 1:
            continue  # edge/corner: several solid voxels share the entry t
        assert got[i]["hit"] == ref[i]["hit"], (i, rays[i], got[i], ref[i])
        if ref[i]["hit"]:
            assert tuple(got[i]["voxel"]) == tuple(ref[i]["voxel"]), (i, got[i], ref[i])
            assert abs(got[i]["t"] - ref[i]["t"]) < 1e-3
            assert got[i]["material"] == ref[i]["material"]
            n_checked += 1
    assert n_checked > 60


def test_r1_matches_r0_on_clean_primary_hits(small_world):
    wl = workloads.SMALL
    _, cam = wl.camera()
    ow = oracle.OracleWorld.view(small_world)
    rows = (300, 340)
    r1, _, _ = ow.frame(cam, wl.width, wl.height, wl.primary_tmax, 0, None, 0.0, rows)
    r0, _ = ow.frame_r0(cam, wl.width, wl.height, wl.primary_tmax, None, 0.0, rows)
    dirs = oracle.primary_dirs(cam, wl.width, wl.height)[rows[0] * wl.width:rows[1] * wl.width]
    eyes = np.repeat(cam[None, :3], len(dirs), 0)
    clean, _, _, grazing, bad = scenes.classify(eyes, dirs, r1, r0)
    assert not bad, bad[:3]
    assert clean > 0.9 * r1["hit"].sum()
    agree = (r1["hit"] == r0["hit"]) & np.all(r1["voxel"] == r0["voxel"], axis=1)
    assert agree.mean() > 0.99
    # the vectorised classifier agrees with the per-ray one it replaces at scale
    c = oracle.classify_vs_r0(cam[:3], dirs, r1, r0, tmax=wl.primary_tmax)
    assert c["agree_clean"] == clean and c["clean_disagree"] == 0


def test_r1_vs_r0_small_frame_full(small_world):
    # the 1280x720 frame of the 256x128x256 world, every pixel
    wl = workloads.SMALL
    _, cam = wl.camera()
    ow = oracle.OracleWorld.view(small_world)
    r1, _, _ = ow.frame(cam, wl.width, wl.height, wl.primary_tmax, 0, None, 0.0)
    r0, _ = ow.frame_r0(cam, wl.width, wl.height, wl.primary_tmax, None, 0.0)
    dirs = oracle.primary_dirs(cam, wl.width, wl.height)
    c = oracle.classify_vs_r0(cam[:3], dirs, r1, r0, tmax=wl.primary_tmax)
    scenes.assert_agrees_with_r0(c, "small frame R1 vs R0")
    assert c["agree_clean"] > 0.85 * r0["hit"].sum()


def test_r1_vs_r0_batched_step_full(small_world):
    # one whole step of the 96 uneven batches (primary, shadow any-hit and AO rays, tmax in
    # {32, 128, 512, 1024}); R0 normalizes every direction itself, as the GPU entry does
    rays, _ = workloads.BATCHED.rays(small_world)
    ow = oracle.OracleWorld.view(small_world)
    r1, _ = ow.trace_rays(rays)
    r0 = ow.raycast_rays(rays)
    d = oracle.normalize_or_zero_v(rays["d"])
    c = oracle.classify_vs_r0(rays["o"], d, r1, r0, tmax=rays["tmax"])
    scenes.assert_agrees_with_r0(c, "batched step R1 vs R0")
    assert c["agree_clean"] > 0.8 * r0["hit"].sum()


def test_r1_vs_r0_large_frame_rows():
    # rows of the 1920x1080 frame of the 2048x256x2048 world with brick / tile LOD past
    # 400 / 520: compared where the LOD frame equals the exact walk (no hit at t >= 400),
    # plus the shadow rays of the agreeing hits
    wl = workloads.LARGE
    world = wl.world()
    _, cam = wl.camera()
    sun = wl.sun()
    ow = oracle.OracleWorld.view(world)
    rows = (300, 340)  # past the horizon's far LOD hits and down to near ground
    r1, s1, _ = ow.frame(cam, wl.width, wl.height, wl.primary_tmax, wl.primary_flags, sun,
                         wl.shadow_tmax, rows)
    r0, s0 = ow.frame_r0(cam, wl.width, wl.height, wl.primary_tmax, sun, wl.shadow_tmax, rows)
    dirs = oracle.primary_dirs(cam, wl.width, wl.height)[rows[0] * wl.width:rows[1] * wl.width]
    sel = (r1["hit"] == 0) | (r1["t"] < 400.0)
    c = oracle.classify_vs_r0(cam[:3], dirs[sel], r1[sel], r0[sel], tmax=wl.primary_tmax)
    scenes.assert_agrees_with_r0(c, "large frame rows R1 vs R0")
    assert c["agree_clean"] > 1000 and (~sel).sum() > 1000
    same_ray = scenes.same_shadow_ray(r1, r0, cam[:3], dirs, sel)
    cs = oracle.classify_shadows_vs_r0(ow, cam[:3], dirs, r1, s1, s0, sun, wl.shadow_tmax,
                                       same_ray)
    print("large frame rows shadows", cs)
    assert cs["clean_disagree"] == 0 and cs["rays"] > 1000, cs


def test_phase1_leaves_solid_start():
    # src/raycast.rs:108-115: a ray starting inside solid exits it before looking for a hit
    ow = oracle.OracleWorld(scenes.REF_DIMS)
    scenes.fill_box(ow, 10, 14, 10, 14, 10, 14, 4)  # 4^3 solid block
    scenes.fill_box(ow, 20, 21, 10, 14, 10, 14, 3)  # wall further along +x
    o, d = np.array([11.5, 11.5, 11.5], np.float32), np.array([1, 0, 0], np.float32)
    r0 = ow.raycast_r0(o, d)
    r1, _ = ow.trace_r1(o, d)
    assert tuple(r0["voxel"]) == tuple(r1["voxel"]) == (20, 11, 11)
    assert r1["material"] == 3 and r1["face"] == 1
    assert r1["t"] == np.float32(20.0 - 11.5)


def test_tmax_and_entry_t():
    ow = oracle.OracleWorld(scenes.REF_DIMS)
    scenes.fill_box(ow, 40, 41, 0, 256, 0, 512, 4)  # wall at x = 40
    o, d = np.array([0.5, 100.5, 100.5], np.float32), np.array([1, 0, 0], np.float32)
    h, _ = ow.trace_r1(o, d, tmax=39.5)
    assert h["hit"] and h["t"] == np.float32(39.5)  # entry t == tmax still hits
    h, _ = ow.trace_r1(o, d, tmax=39.49)
    assert not h["hit"]
    # rays from outside the window enter at the slab entry
    o2 = np.array([-100.0, 100.5, 100.5], np.float32)
    h, _ = ow.trace_r1(o2, d)
    assert h["hit"] and tuple(h["voxel"]) == (40, 100, 100) and h["t"] == np.float32(140.0)
    # zero direction misses
    h, _ = ow.trace_r1(o, np.zeros(3, np.float32))
    assert not h["hit"] and h["t"] == -1.0


def test_lod_topmost_material():
    # brick LOD past 400 reports the brick's topmost non-turf material at the ray's entry
    # voxel; tile LOD past 520 the topmost of the tile's first occupied brick
    # (shaders/raymarch.wgsl:253-295, 4673-4758, 4782-4798)
    ow = oracle.OracleWorld(scenes.REF_DIMS)
    # brick at bricks (110..111) along x at y 100..103, z 100..103: layers y=0..3
    ow.set_voxel(441, 100, 101, 4)      # y layer 0 stone
    ow.set_voxel(442, 102, 101, 3)      # y layer 2 dirt
    ow.set_voxel(440, 103, 102, 35)     # y layer 3, lowest set bit: turf
    ow.set_voxel(443, 103, 103, 2)      # y layer 3, higher bit: grass
    o = np.array([0.5, 101.5, 100.5], np.float32)
    d = np.array([1, 0, 0], np.float32)
    h, _ = ow.trace_r1(o, d, flags=oracle.LOD)
    # each layer is decided by its lowest set bit only; a turf there drops to the next layer,
    # so the answer is layer 2's dirt, not layer 3's grass
    assert h["hit"] and tuple(h["voxel"]) == (440, 101, 100) and h["material"] == 3
    assert h["t"] == np.float32(439.5) and h["face"] == 1
    h, _ = ow.trace_r1(o, d)  # LOD off: the true first solid voxel along y=101,z=100 row
    assert not h["hit"]
    # tile LOD: far tile whose first occupied brick (lowest bit) has a grass top
    ow2 = oracle.OracleWorld(scenes.REF_DIMS)
    ow2.set_voxel(496, 96, 96, 4)
    ow2.set_voxel(496, 97, 96, 2)
    ow2.set_voxel(505, 100, 100, 3)
    o = np.array([-30.0, 100.5, 100.5], np.float32)  # tile entry at t = 526 > 520
    h, c = ow2.trace_r1(o, d, flags=oracle.LOD)
    assert h["hit"] and h["material"] == 2 and tuple(h["voxel"]) == (496, 100, 100)
    assert h["t"] == np.float32(526.0) and c["tile_lod"] == 1
    o = np.array([0.5, 100.5, 100.5], np.float32)  # tile entered at t = 495.5: brick LOD
    h, c = ow2.trace_r1(o, d, flags=oracle.LOD)
    assert h["hit"] and h["material"] == 3 and tuple(h["voxel"]) == (504, 100, 100)
    assert c["tile_lod"] == 0 and c["brick_lod"] == 1


@pytest.mark.parametrize("flags", [0, oracle.LOD])
def test_group_skips_keep_results(small_world, flags):
    # 2x2x2 empty-group skips (and their LOD guard) change how the walk moves, never
    # what it returns: same frame bit for bit with and without them
    wl = workloads.SMALL
    _, cam = wl.camera()
    sun = workloads.LARGE.sun()
    ow = oracle.OracleWorld.view(small_world)
    # (the y-bound exit off, so the walks above the terrain skip groups too)
    a, sa, ca = ow.frame(cam, wl.width, wl.height, 700.0, flags | oracle.NO_YTOP, sun, 480.0)
    b, sb, cb = ow.frame(cam, wl.width, wl.height, 700.0,
                         flags | oracle.NO_GROUPS | oracle.NO_YTOP, sun, 480.0)
    assert a.tobytes() == b.tobytes() and np.array_equal(sa, sb)
    assert cb["brick_group_skip"] == 0 and cb["tile_group_skip"] == 0
    assert ca["brick_group_skip"] > 0 and ca["tile_group_skip"] > 0
    assert sum(ca.values()) < sum(cb.values())


@pytest.mark.parametrize("wl", [workloads.SMALL, workloads.LARGE], ids=["small", "large"])
def test_ytop_exit_keeps_results(wl):
    # a walk at y >= ytop that does not move down ends with a miss: the same frame (primary
    # LOD rays and their shadow rays) as without the exit, in fewer iterations; the bound is
    # the highest occupied brick row's top
    world = workloads.SMALL.world() if wl is workloads.SMALL else wl.world()
    _, cam = wl.camera()
    sun = workloads.LARGE.sun()
    ow = oracle.OracleWorld.view(world)
    top = oracle.world_ytop(world.dims, world.origin, world.tile_mask, world.tile_uniform)
    occ = world.bricks["occupancy"].reshape(world.brick_dims[::-1]) != 0
    rows = np.nonzero(occ.any(axis=(0, 2)))[0]
    assert top == 4 * (int(rows.max()) + 1) < world.dims[1]
    rows_traced = None if wl is workloads.SMALL else (0, 1080, 4)
    a, sa, ca = ow.frame(cam, wl.width, wl.height, wl.primary_tmax, oracle.LOD, sun, 480.0,
                         rows_traced)
    b, sb, cb = ow.frame(cam, wl.width, wl.height, wl.primary_tmax, oracle.LOD | oracle.NO_YTOP,
                         sun, 480.0, rows_traced)
    assert a.tobytes() == b.tobytes() and np.array_equal(sa, sb)
    assert sum(ca.values()) < 0.95 * sum(cb.values())


def test_group_skips_keep_lod_results_far():
    # far LOD region of the large world: every group-skip guard boundary is exercised
    wl = workloads.LARGE
    world = wl.world()
    _, cam = wl.camera()
    sun = wl.sun()
    ow = oracle.OracleWorld.view(world)
    rows = (240, 330)  # rows around the horizon reach past t 400 / 520
    a, sa, ca = ow.frame(cam, wl.width, wl.height, 700.0, oracle.LOD, sun, 480.0, rows)
    b, sb, _ = ow.frame(cam, wl.width, wl.height, 700.0, oracle.LOD | oracle.NO_GROUPS, sun,
                        480.0, rows)
    assert a.tobytes() == b.tobytes() and np.array_equal(sa, sb)
    assert ca["tile_lod"] > 0 and ca["brick_lod"] > 0
