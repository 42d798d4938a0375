"""On-device terrain generation (vg_world_generate) leaves exactly the world the host
generator + pyramid builder + upload leave: bricks, masks and uniform tables, with and
without a folded window origin."""
import numpy as np
import pytest

import voxelg_amd as vg
from voxelg_amd import workloads

pytestmark = pytest.mark.gpu

FIELDS = ("bricks", "tile_mask", "chunk_mask", "l4_mask", "brick_uniform", "tile_uniform")


@pytest.mark.parametrize("dims,origin,seed", [
    (workloads.SMALL.dims, (0, 0, 0), workloads.SMALL.seed),
    ((512, 256, 512), (320, 0, -192), 0x1234_5678_9ABC_DEF0),
    ((256, 64, 512), (0, 0, 0), 7),
])
def test_generate_matches_host(ctx, dims, origin, seed):
    print(f"case: generate {dims} origin {origin} seed {seed:#x}", flush=True)
    host = vg.World(dims, origin).generate(seed)
    got = ctx.world(dims).generate(seed, origin).download()
    for f in FIELDS:
        a, b = getattr(got, f), getattr(host, f)
        bad = np.nonzero(a != b)[0]
        assert len(bad) == 0, f"{f} differs at {bad[:5]} ({len(bad)} entries)"
    assert host.solid_fraction() > 0.05
