"""Pin the oracle (R0 = src/raycast.rs restated; R1 = the exact hierarchical walk the GPU
must match) on the reference's own known-answer scenes, src/accel.rs:515-942."""
import numpy as np
import pytest

import oracle
import scenes


def r0_hits(ow, eyes, dirs, tmax=np.inf):
    return np.array([ow.raycast_r0(e, d, tmax) for e, d in zip(eyes, dirs)])


def r1_hits(ow, eyes, dirs, tmax=np.inf, flags=0):
    out, _ = ow.trace_rays(oracle.make_rays(eyes, dirs, tmax), flags)
    return out


@pytest.mark.parametrize("impl", ["r0", "r1"])
def test_flat_floor_surface_height(impl):
    # src/accel.rs:528-572: near-vertical rays onto the y=60 layer enter its top face y=61
    ow = scenes.flat_floor(16, 16, 48)
    eyes, dirs = [], []
    for gz in range(6):
        for gx in range(6):
            eyes.append([22.0 + gx * 6.0, 120.0, 22.0 + gz * 6.0])
            d = np.array([0.03, -1.0, 0.017], np.float32)
            dirs.append(d / np.float32(np.linalg.norm(d)))
    eyes, dirs = np.array(eyes, np.float32), np.array(dirs, np.float32)
    hits = r0_hits(ow, eyes, dirs) if impl == "r0" else r1_hits(ow, eyes, dirs)
    for i, h in enumerate(hits):
        assert h["hit"] == 1, i
        assert h["voxel"][1] == 60, (i, h)
        hit_y = eyes[i][1] + h["t"] * dirs[i][1]
        assert abs(hit_y - 61.0) < 0.05, (i, hit_y)
        assert h["face"] == 2  # entered through the +y face
        assert h["material"] == scenes.STONE


@pytest.mark.parametrize("origin", [(0, 0, 0), (96, 0, 160)])
def test_floor_pillars_r1_matches_r0(origin):
    # src/accel.rs:655-759: clean hits must agree exactly (voxel and normal); edge/corner
    # grazing hits (lattice_touch >= 2) are classified, never asserted.
    ow = scenes.floor_pillars(origin)
    eyes, dirs = scenes.floor_pillar_rays(origin)
    ref = r0_hits(ow, eyes, dirs)
    got = r1_hits(ow, eyes, dirs)
    clean, floor, pillar, grazing, bad = scenes.classify(eyes, dirs, got, ref)
    total = len(eyes)
    assert not bad, bad[:4]
    assert clean > total / 3, (clean, total)
    assert floor > 0 and pillar > 0
    assert grazing < clean


def test_floor_pillars_hits_are_world_coords_under_shift():
    # the shifted window (origin chunk (3,5) = voxel (96,0,160)) must report WORLD voxels
    a = scenes.floor_pillars((0, 0, 0))
    b = scenes.floor_pillars((96, 0, 160))
    e0, d0 = scenes.floor_pillar_rays((0, 0, 0))
    e1, d1 = scenes.floor_pillar_rays((96, 0, 160))
    h0 = r1_hits(a, e0, d0)
    h1 = r1_hits(b, e1, d1)
    assert np.array_equal(h0["hit"], h1["hit"])
    m = h0["hit"] == 1
    assert np.array_equal(h0["voxel"][m] + np.array([96, 0, 160]), h1["voxel"][m])


def test_grow_shrink_rebuild():
    # src/accel.rs:761-834: a pillar raised in place is hit; removed again, the ray misses
    ow = scenes.flat_floor(8, 8, 56)
    scenes.fill_box(ow, 30, 33, 61, 75, 30, 33, scenes.STONE)
    eye, d = np.array([20.0, 70.0, 31.5], np.float32), np.array([1.0, 0.0, 0.0], np.float32)
    r0 = ow.raycast_r0(eye, d)
    r1, _ = ow.trace_r1(eye, d)
    assert r0["hit"] and r1["hit"]
    assert tuple(r0["voxel"]) == tuple(r1["voxel"]) == (30, 70, 31)
    scenes.fill_box(ow, 30, 33, 61, 75, 30, 33, 0)
    assert not ow.raycast_r0(eye, d)["hit"]
    assert not ow.trace_r1(eye, d)[0]["hit"]
    # masks after the incremental edits equal a full rebuild
    ref = oracle.OracleWorld(ow.dims)
    ref.bricks[:] = ow.bricks
    ref.rebuild()
    for a in ("tile_mask", "chunk_mask", "l4_mask", "brick_uniform", "tile_uniform"):
        assert np.array_equal(getattr(ow, a), getattr(ref, a)), a


def test_roof_shadow_occlusion():
    # src/accel.rs:836-917: any-hit toward the sun; decisive samples agree, both classes occur
    ow = scenes.roof_scene()
    origins, sun, under, open_ = scenes.roof_samples()
    dirs = np.repeat(sun[None], len(origins), 0)
    cpu = r0_hits(ow, origins, dirs)["hit"] == 1
    got = r1_hits(ow, origins, dirs, flags=oracle.ANY_HIT)["hit"] == 1
    decisive = under | open_
    assert np.sum((cpu != got) & decisive) == 0
    assert np.sum(under & got) > 20 and np.sum(open_ & ~got) > 20
    # any-hit == (first hit exists within tmax), here unbounded
    first = r1_hits(ow, origins, dirs)["hit"] == 1
    assert np.array_equal(first, got)


def test_empty_world_all_miss():
    # src/accel.rs:919-942
    ow = oracle.OracleWorld(scenes.REF_DIMS)
    eyes = np.array([[20 + gx * 30.0, 200.0, 20 + gz * 30.0] for gz in range(8) for gx in range(8)],
                    np.float32)
    dirs = np.repeat(np.array([[0.05, -1.0, 0.03]], np.float32), 64, 0)
    assert not np.any(r0_hits(ow, eyes, dirs)["hit"])
    hits, counts = ow.trace_rays(oracle.make_rays(eyes, dirs))
    assert not np.any(hits["hit"])
    assert np.all(hits["t"] == -1.0) and np.all(hits["face"] == oracle.FACE_NONE)
    # an empty world is crossed in whole-L4 skips only
    assert counts["voxel"] == 0 and counts["l4_skip"] > 0
