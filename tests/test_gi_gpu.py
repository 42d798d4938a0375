"""GPU parity of the probe update (vg_probes_update) against the FP64 oracle on the same
samples: every gather / shadow / bounce ray is traced bit-exactly on both sides, so the
records agree to fp32 summation error in the SH sums and exactly in the integer-valued
fields (grid cell, age, staging count) and the fp32 relocation offsets."""
import math
import os

import numpy as np
import pytest

import oracle
from voxelg_amd import gi
from voxelg_amd._lib import VGError

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
SUN = (math.cos(math.radians(35.0)) * math.cos(2.1), math.sin(math.radians(35.0)),
       math.cos(math.radians(35.0)) * math.sin(2.1))


def as_fields(probes: np.ndarray) -> np.ndarray:
    return probes.view(np.float32).reshape(len(probes), 48).astype(np.float64)


def assert_records_match(got: np.ndarray, ref: np.ndarray, what: str):
    g = as_fields(got)
    # integer-valued and fp32 fields: exact (info, staging count, relocation offset)
    for col in (20, 21, 22, 23, 44, 45, 46, 47):
        bad = np.nonzero(g[:, col] != ref[:, col])[0]
        assert len(bad) == 0, f"{what}: field {col} differs at {bad[:4]}: {g[bad[:2], col]} " \
                              f"vs {ref[bad[:2], col]}"
    # SH sums (displayed 0..19, staging 24..43): fp32 summation of the same samples
    for base in (0, 24):
        for grp in range(5):
            cols = slice(base + 4 * grp, base + 4 * grp + 4)
            scale = max(np.abs(ref[:, base + 4 * grp]).max(), 1e-6)
            err = np.abs(g[:, cols] - ref[:, cols])
            tol = 2e-5 * scale + 2e-5 * np.abs(ref[:, cols])
            bad = np.argwhere(err > tol)
            assert len(bad) == 0, f"{what}: SH group {base}+{grp} off at {bad[:3]}: " \
                                  f"{g[bad[0][0], cols]} vs {ref[bad[0][0], cols]}"


def run_rounds(ctx, dw, world, params, rounds, check_every=1):
    dirs = gi.directions(int(params[0]["rays_per_probe"]) * int(params[0]["epochs"]))
    bt, pal = gi.bounce_table(), gi.default_palette()
    grid = gi.ProbeGrid(dw, params)
    ow = oracle.OracleWorld.view(world)
    ref = np.zeros((grid.total, 48), np.float64)
    p = params.copy()
    for r in rounds:
        p["round"] = r
        print(f"case: probes {tuple(p[0]['grid'])} rays {p[0]['rays_per_probe']} "
              f"bounces {p[0]['bounces']} round {r}", flush=True)
        grid.update(p)
        ow.gi_update(p, dirs, bt, pal, ref, threads=THREADS)
        if (r - rounds[0]) % check_every == 0 or r == rounds[-1]:
            ctx.synchronize()
            assert_records_match(grid.download(), ref, f"round {r}")
    return grid, ref


def test_plain_update_on_a_folded_window(ctx):
    # a window whose origin folds x and z into toroidal slots runs the path kernel's folding
    # instantiation: the records against the oracle's on the same window
    import voxelg_amd as vg
    from voxelg_amd import workloads
    world = vg.World(workloads.SMALL.dims, origin=(96, 0, 160)).generate(workloads.SMALL.seed)
    dw = ctx.world(world.dims).upload(world)
    p = gi.gi_params(SUN, grid=(8, 4, 8), spacing=32, rays_per_probe=32, epochs=1,
                     update_div=1, bounces=2)
    _, ref = run_rounds(ctx, dw, world, p, range(2))
    assert (ref[:, 23] == 1.0).sum() > 50
    dw.close()


def test_plain_two_bounce_matches_oracle(ctx, small, small_world):
    p = gi.gi_params(SUN, grid=(8, 4, 8), spacing=32, rays_per_probe=128, epochs=2,
                     update_div=1, bounces=2)
    grid, ref = run_rounds(ctx, small, small_world, p, range(4))
    live = ref[:, 23] == 1.0
    assert 0.3 < live.mean() < 1.0  # some probes buried in the terrain, most live
    assert np.any(ref[live, 44] != 0) or np.any(ref[live, :20] != 0)


def test_reference_defaults_one_bounce_matches_oracle(ctx, small, small_world):
    """The reference's own schedule: 8 rays, 8 direction epochs, 1/8 of the probes per
    round, one bounce, spacing 8 (a 32 x 16 x 32 grid over the 256 x 128 x 256 world), over
    72 rounds: every probe seeds, stages all 8 epochs, folds once and relocates."""
    p = gi.gi_params(SUN, grid=(32, 16, 32))
    grid, ref = run_rounds(ctx, small, small_world, p, range(72), check_every=8)
    live = ref[:, 23] == 1.0
    assert live.mean() > 0.3
    assert np.count_nonzero(ref[live][:, 45:48]) > 0  # some probes relocated


@pytest.mark.parametrize("grid,rays,div,bounces", [((2, 1, 2), 8, 1, 2), ((4, 2, 4), 1, 3, 1),
                                                    ((2, 2, 2), 24, 4, 2), ((3, 1, 3), 5, 2, 1)])
def test_small_updates_match_oracle(ctx, small, small_world, grid, rays, div, bounces):
    # updates of fewer probes than the path kernel's 8 work stripes (empty stripes, ragged
    # last stripe), odd ray counts and partial updates (update_div > 1) over several rounds
    p = gi.gi_params(SUN, grid=grid, spacing=32, rays_per_probe=rays, epochs=1,
                     update_div=div, bounces=bounces)
    run_rounds(ctx, small, small_world, p, range(2 * div))


def test_ray_counts_match_oracle(ctx, small, small_world):
    p = gi.gi_params(SUN, grid=(8, 4, 8), spacing=32, rays_per_probe=64, epochs=1,
                     update_div=1, bounces=2)
    grid = gi.ProbeGrid(small, p)
    rays, cnt = grid.update_counts(p)
    ref = np.zeros((grid.total, 48))
    orays = oracle.OracleWorld.view(small_world).gi_update(
        p, gi.directions(64), gi.bounce_table(), gi.default_palette(), ref, threads=THREADS)
    assert [rays["gather"], rays["shadow"], rays["bounce"]] == orays
    assert rays["bounce"] > 0 and rays["shadow"] > 0
    assert sum(cnt.values()) > rays["gather"]
    assert_records_match(grid.download(), ref, "counts run")


def test_palette_and_reset(ctx, small, small_world):
    p = gi.gi_params(SUN, grid=(4, 4, 4), spacing=64, rays_per_probe=64, epochs=1,
                     update_div=1, bounces=1)
    pal = gi.default_palette()
    pal[1:40, :3] *= 0.5
    grid = gi.ProbeGrid(small, p).set_palette(pal)
    grid.update(p)
    ref = np.zeros((grid.total, 48))
    oracle.OracleWorld.view(small_world).gi_update(p, gi.directions(64), gi.bounce_table(), pal,
                                                   ref, threads=THREADS)
    ctx.synchronize()
    assert_records_match(grid.download(), ref, "palette")
    grid.reset()
    ctx.synchronize()
    assert np.all(as_fields(grid.download()) == 0)


def test_param_errors(ctx, small):
    p = gi.gi_params(SUN, grid=(4, 4, 4), spacing=64, rays_per_probe=16, epochs=1)
    grid = gi.ProbeGrid(small, p)
    q = p.copy()
    q["grid"] = (4, 4, 8)
    with pytest.raises(VGError):
        grid.update(q)
    q = p.copy()
    q["rays_per_probe"] = 32
    with pytest.raises(VGError):
        grid.update(q)
    q = p.copy()
    q["bounces"] = 3
    with pytest.raises(VGError):
        grid.update(q)


@pytest.mark.parametrize("bounces", [1, 2])
def test_schedules_give_the_same_records(ctx, small, small_world, bounces):
    """The probe path kernel's schedule (waves per CU, refill threshold, work ids per atomic,
    transition batch, material prefetch) and its longest-first order (each update traces the previous update's heavy paths first
    and skips them in the natural order) change only the order of the work: every schedule
    gives the oracle's records, round after round."""
    from voxelg_amd import _lib as L
    opts = (L.VG_OPT_PROBE_WAVES_PER_CU, L.VG_OPT_PROBE_REFILL, L.VG_OPT_PROBE_CHUNK,
            L.VG_OPT_PROBE_TBATCH, L.VG_OPT_PROBE_MATSPEC)
    keep = [ctx.get_option(o) for o in opts]
    try:
        for wpc, refill, chunk, tb, ms in ((8, 4, 0, 1, 0), (16, 48, 16, 8, 1), (4, 60, 128, 64, 0),
                                           (24, 32, 1024, 16, 1), (8, 1, 32, 64, 1)):
            for o, v in zip(opts, (wpc, refill, chunk, tb, ms)):
                ctx.set_option(o, v)
            print(f"case: schedule {wpc} waves/CU, refill {refill}, chunk {chunk}, "
                  f"transition batch {tb}, material prefetch {ms}", flush=True)
            p = gi.gi_params(SUN, grid=(8, 4, 8), spacing=32, rays_per_probe=64, epochs=1,
                             update_div=1, bounces=bounces)
            run_rounds(ctx, small, small_world, p, range(3))
    finally:
        for o, v in zip(opts, keep):
            ctx.set_option(o, v)


def test_schedule_option_errors(ctx):
    from voxelg_amd import _lib as L
    for bad in (-1, 1, 8, 15, 1025):
        with pytest.raises(VGError):
            ctx.set_option(L.VG_OPT_PROBE_CHUNK, bad)
    for bad in (0, -3, 65):
        with pytest.raises(VGError):
            ctx.set_option(L.VG_OPT_PROBE_TBATCH, bad)
    for bad in (-1, 2):
        with pytest.raises(VGError):
            ctx.set_option(L.VG_OPT_PROBE_MATSPEC, bad)
    keep = ctx.get_option(L.VG_OPT_PROBE_CHUNK)
    ctx.set_option(L.VG_OPT_PROBE_CHUNK, 256)
    assert ctx.get_option(L.VG_OPT_PROBE_CHUNK) == 256
    ctx.set_option(L.VG_OPT_PROBE_CHUNK, keep)


def indirect_case(small, p, cam, W, H, pg):
    hits, _ = small.trace_frame(cam, W, H, 512.0, 0, SUN, 480.0)
    out = pg.frame_indirect(p, cam, W, H, hits)
    got = out.numpy(np.float32).reshape(-1, 4)
    grid = tuple(int(v) for v in p[0]["grid"])
    ref = oracle.frame_indirect(pg.download(), grid, int(p[0]["spacing"]), cam, W, H,
                                hits.numpy(np.uint8))
    return got, ref


def test_frame_indirect_matches_oracle(ctx, small, small_world):
    """The probe cache's per-pixel consumer (indirect_light over sample_probes) equals the
    oracle's fp32 restatement bit for bit: on records of real updates, and on random records
    (relocation offsets, distance moments that trigger the Chebyshev bound, ungathered and
    foreign slots)."""
    import voxelg_amd as vg
    p = gi.gi_params(SUN, grid=(8, 4, 8), spacing=32, rays_per_probe=32, epochs=1,
                     update_div=1, bounces=2, flags=1)
    pg = gi.ProbeGrid(small, p)
    for r in (0, 1):
        p["round"] = r
        pg.update(p)
    W, H = 192, 108
    cam = vg.Camera(pos=(128.5, 90.0, 128.5), yaw=0.7, pitch=-0.35).uniform(W, H)
    print("case: frame indirect over updated probes", flush=True)
    got, ref = indirect_case(small, p, cam, W, H, pg)
    assert got.tobytes() == ref.tobytes()
    assert (got[:, :3] > 0).mean() > 0.2
    # random records
    rng = np.random.default_rng(21)
    rec = np.zeros(pg.total, gi.PROBE_DTYPE)
    f = rec.view(np.float32).reshape(pg.total, 48)
    f[:, 0:12] = rng.normal(0.5, 0.6, (pg.total, 12))
    f[:, 12] = rng.uniform(0.0, 200.0, pg.total)          # mean distance DC
    f[:, 13:16] = rng.normal(0.0, 20.0, (pg.total, 3))
    f[:, 16] = f[:, 12] ** 2 * rng.uniform(0.5, 1.5, pg.total)
    f[:, 17:20] = rng.normal(0.0, 100.0, (pg.total, 3))
    gx, gy, gz = p[0]["grid"]
    idx = np.arange(pg.total)
    f[:, 20], f[:, 21], f[:, 22] = idx % gx, (idx // gx) % gy, idx // (gx * gy)
    f[:, 23] = rng.choice((0.0, 1.0, 5.0), pg.total, p=(0.1, 0.3, 0.6))
    f[rng.random(pg.total) < 0.05, 20] += gx             # a foreign world cell
    f[:, 45:48] = rng.uniform(-6.0, 6.0, (pg.total, 3))   # relocation offsets
    pg.upload(rec)
    for cam in (vg.Camera(pos=(128.5, 90.0, 128.5), yaw=0.7, pitch=-0.35).uniform(W, H),
                vg.Camera(pos=(40.0, 75.0, 200.0), yaw=2.4, pitch=-0.2).uniform(W, H)):
        print("case: frame indirect over random records", flush=True)
        got, ref = indirect_case(small, p, cam, W, H, pg)
        assert got.tobytes() == ref.tobytes()
        assert (got[:, :3] > 0).mean() > 0.1
    q = p.copy()
    q["grid"] = (4, 4, 4)
    hits, _ = small.trace_frame(cam, W, H, 512.0, 0, SUN, 480.0)
    with pytest.raises(VGError):
        pg.frame_indirect(q, cam, W, H, hits)
    # the pixel rays come from the camera: a resolution other than the frame's is refused
    bad = vg.Camera(pos=(128.5, 90.0, 128.5), yaw=0.7, pitch=-0.35).uniform(W // 2, H)
    with pytest.raises(VGError):
        pg.frame_indirect(p, bad, W, H, hits)
