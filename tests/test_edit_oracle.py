"""Pins of the oracle's sphere-edit replay (orc_apply_spheres): it equals the reference's
per-voxel path (apply_edit -> set_voxel with its incremental mask / uniform refresh,
src/voxel.rs:450-474, 871-946), voxel by voxel in sphere order, including windows with a
folded (toroidal) origin and spheres that leave the window."""
import numpy as np
import pytest

import oracle


def spheres(rng, n, dims, origin, r_hi=9):
    e = np.zeros(n, oracle.EDIT_DTYPE)
    lo = np.array(origin) - 6
    hi = np.array(origin) + np.array(dims) + 6  # some spheres leave the window
    e["center"] = rng.integers(lo, hi, (n, 3)).astype(np.int32)
    e["radius"] = rng.integers(0, r_hi + 1, n)
    e["material"] = np.where(rng.random(n) < 0.5, 0, rng.integers(1, 13, n))
    return e


def per_voxel(w, edits):
    for e in edits:
        r = int(e["radius"])
        cx, cy, cz = (int(v) for v in e["center"])
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                for dz in range(-r, r + 1):
                    if dx * dx + dy * dy + dz * dz <= r * r:
                        w.apply_edit(cx + dx, cy + dy, cz + dz, int(e["material"]))


def assert_same(a, b):
    for f in ("bricks", "tile_mask", "chunk_mask", "l4_mask", "brick_uniform", "tile_uniform"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


@pytest.mark.parametrize("origin", [(0, 0, 0), (320, 0, -192)])
def test_apply_spheres_equals_per_voxel_set_voxel(origin):
    dims = (256, 128, 256)
    base = oracle.OracleWorld(dims, origin).generate(0x5EED_0001)
    base.rebuild()
    rng = np.random.default_rng(21)
    edits = spheres(rng, 40, dims, origin)
    # a brick carved empty and a whole tile filled: mask bits and uniform tables change
    extra = np.zeros(2, oracle.EDIT_DTYPE)
    extra[0] = ((origin[0] + 40, 60, origin[2] + 40), 9, 0, (0, 0))
    extra[1] = ((origin[0] + 100, 20, origin[2] + 100), 14, 4, (0, 0))
    edits = np.concatenate([edits, extra])
    a = oracle.OracleWorld.copy_of(base)
    a.apply_spheres(edits)
    b = oracle.OracleWorld.copy_of(base)
    per_voxel(b, edits)
    assert_same(a, b)
    assert not np.array_equal(a.bricks, base.bricks)
    # the filled sphere made at least one uniform brick of stone
    assert np.any(a.brick_uniform == 4)
