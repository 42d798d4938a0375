"""GPU parity of the dynamic world: sphere edits applied on the device (vg_world_apply_edits)
leave exactly the world the oracle's sequential replay + full rebuild leaves (bricks, masks,
uniform tables), and a frame retraced over the dirtied 8x8 blocks only
(vg_frame_dirty_blocks + vg_trace_frame_blocks) equals the oracle's full frame of the edited
world bit for bit, hits and shadow bytes."""
import math

import numpy as np
import pytest

import oracle
import voxelg_amd as vg
from voxelg_amd import workloads
from voxelg_amd._lib import EDIT_DTYPE, VGError
from voxelg_amd.device import decode_hits

pytestmark = pytest.mark.gpu

FIELDS = ("bricks", "tile_mask", "chunk_mask", "l4_mask", "brick_uniform", "tile_uniform")


def case(msg):
    print(f"case: {msg}", flush=True)


def assert_world_equal(got, ref, what):
    for f in FIELDS:
        a, b = getattr(got, f), getattr(ref, f)
        bad = np.nonzero(a != b)[0]
        assert len(bad) == 0, f"{what}: {f} differs at {bad[:5]} ({len(bad)} entries)"


def edge_edits(dims, origin):
    """Spheres across the window faces, outside it, a single voxel, a carved brick and a
    filled tile."""
    o = np.array(origin)
    d = np.array(dims)
    e = np.zeros(7, EDIT_DTYPE)
    e[0] = (tuple(o + [0, 40, 30]), 6, 7, (0, 0))            # straddles x = window start
    e[1] = (tuple(o + d - [1, 1, 1]), 5, 0, (0, 0))           # window corner
    e[2] = (tuple(o + [-20, 50, 50]), 4, 3, (0, 0))           # entirely outside
    e[3] = (tuple(o + [33, 70, 91]), 0, 9, (0, 0))            # one voxel
    e[4] = (tuple(o + [101, 62, 45]), 8, 0, (0, 0))           # carve
    e[5] = (tuple(o + [150, 30, 150]), 14, 4, (0, 0))         # fill (uniform bricks / tile)
    e[6] = (tuple(o + [150, 30, 150]), 3, 0, (0, 0))          # later carve inside the fill
    return e


@pytest.mark.parametrize("origin", [(0, 0, 0), (320, 0, -192)])
def test_edits_match_oracle_world(ctx, origin):
    wl = workloads.DYNAMIC
    world = vg.World(wl.dims, origin).generate(0x5EED_0001)
    ref = oracle.OracleWorld.copy_of(world)
    dw = ctx.world(wl.dims).upload(world)
    for f, n in enumerate((0, 300, 2048)):
        if n == 0:
            edits = edge_edits(wl.dims, origin)
        else:
            edits = wl.edits(f, world, n)
            edits["center"] += np.array(origin, np.int32)
        case(f"edits origin {origin} batch {f}: {len(edits)}")
        dw.apply_edits(edits)
        ref.apply_spheres(edits)
        assert_world_equal(dw.download(), ref, f"batch {f}")
    assert not np.array_equal(ref.bricks, world.bricks)


@pytest.mark.parametrize("flags, pos, yaw, tmax", [
    (0, (128.5, 90.0, 128.5), 0.7, 512.0),
    # the beam prologue over the block list (beam t indexed by block id)
    (vg.VG_RAY_LOD | vg.VG_RAY_BEAM, (128.5, 90.0, 128.5), 0.7, 700.0),
    # brick / tile LOD past t 400 / 520: a camera outside the 256-voxel window looking in
    (vg.VG_RAY_LOD, (-150.5, 110.0, -140.5), 0.393, 700.0)])
def test_dirty_retrace_is_bit_exact(ctx, flags, pos, yaw, tmax):
    wl = workloads.DYNAMIC
    world = wl.world()
    ref = oracle.OracleWorld.copy_of(world)
    dw = ctx.world(wl.dims).upload(world)
    W, H = 640, 360
    cam = vg.Camera(pos=pos, yaw=yaw, pitch=-0.35).uniform(W, H)
    sun = vg.sun_direction(math.radians(35.0), 2.1)
    hits, occ = dw.trace_frame(cam, W, H, tmax, flags, sun, 480.0)
    if flags & vg.VG_RAY_LOD and pos[0] < 0:
        _, _, c0 = ref.frame(cam, W, H, tmax, flags, sun, 480.0)
        assert c0["brick_lod"] > 0 or c0["tile_lod"] > 0, c0
    nbx, nby = (W + 7) // 8, (H + 7) // 8
    rng = np.random.default_rng(5)
    for f, n in enumerate((1, 16, 4096)):
        if n == 1:  # a small carve at a voxel the frame sees: a partial mask
            hh = decode_hits(hits).reshape(H, W)
            ys, xs = np.nonzero(hh["hit"] == 1)
            k = np.argmin(np.abs(ys - H // 2) + np.abs(xs - (3 * W) // 4))
            h0 = hh[ys[k], xs[k]]
            edits = np.zeros(1, EDIT_DTYPE)
            edits[0] = (tuple(int(v) for v in h0["voxel"]), 2, 0, (0, 0))
        else:
            edits = wl.edits(100 + f, world, n)
        before = decode_hits(hits).copy()
        case(f"dirty retrace frame {f}: {n} edits")
        dw.apply_edits(edits)
        ref.apply_spheres(edits)
        mask = dw.dirty_blocks(cam, W, H, sun, 480.0)
        dw.trace_frame_blocks(cam, W, H, tmax, flags, sun, 480.0, mask, hits, occ)
        ctx.synchronize()
        got = decode_hits(hits)
        exp, exp_occ, _ = ref.frame(cam, W, H, tmax, flags, sun, 480.0)
        assert got.tobytes() == exp.tobytes(), f"frame {f}: hits differ"
        assert np.array_equal(occ.numpy(np.uint8), exp_occ), f"frame {f}: shadows differ"
        m = mask.numpy(np.uint32)
        if n == 1:
            bits = np.unpackbits(m[:-1].view(np.uint8), bitorder="little")[:nbx * nby]
            assert m[-1] == 0 and 0 < bits.sum() < nbx * nby, "expected a partial mask"
            # blocks outside the mask kept their records
            blk = (np.arange(H)[:, None] // 8) * nbx + (np.arange(W)[None, :] // 8)
            keep = bits[blk.ravel()] == 0
            assert np.array_equal(got[keep].tobytes(), before[keep].tobytes())


def test_edit_errors(ctx):
    dw = ctx.world((256, 128, 256)).upload(workloads.DYNAMIC.world())
    e = np.zeros(1, EDIT_DTYPE)
    e[0] = ((10, 10, 10), 65, 1, (0, 0))
    with pytest.raises(VGError):
        dw.apply_edits(e)
    dw.apply_edits(np.zeros(0, EDIT_DTYPE))  # an empty batch is fine
