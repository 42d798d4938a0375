"""GPU parity: every result of the HIP path equals the oracle's R1 bit for bit (voxel, t bits,
face, material, hit flag, shadow byte), and GPU-vs-R0 disagreements are edge ties only."""
import os

import numpy as np
import pytest

import oracle
import scenes
import voxelg_amd as vg
from voxelg_amd import workloads

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def gpu_hits(h):
    from voxelg_amd.device import decode_hits
    return decode_hits(h)


def assert_same(got, ref, what):
    gb = got.view(np.uint8).reshape(len(got), -1)
    rb = ref.view(np.uint8).reshape(len(ref), -1)
    bad = np.nonzero((gb != rb).any(1))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(got)} differ, first {bad[:4]}: " \
                          f"gpu {got[bad[:2]]} oracle {ref[bad[:2]]}"


def test_small_primary_frame_bit_exact(ctx, small, small_world):
    wl = workloads.SMALL
    _, cam = wl.camera()
    print(f"case: small primary {wl.width}x{wl.height}", flush=True)
    hits = small.trace_primary(cam, wl.width, wl.height, wl.primary_tmax, wl.primary_flags)
    ctx.synchronize()
    ref, _, _ = oracle.OracleWorld.view(small_world).frame(
        cam, wl.width, wl.height, wl.primary_tmax, wl.primary_flags, None, 0.0, threads=THREADS)
    got = gpu_hits(hits)
    assert_same(got, ref, "small primary")
    assert 0.3 < got["hit"].mean() < 1.0


def test_small_frame_lod_shadow_bit_exact(ctx, small, small_world):
    wl = workloads.SMALL
    _, cam = wl.camera()
    sun = workloads.LARGE.sun()
    print("case: small frame LOD+shadow", flush=True)
    hits, occ = small.trace_frame(cam, wl.width, wl.height, 700.0, vg.VG_RAY_LOD, sun, 480.0)
    ctx.synchronize()
    ref, ref_occ, _ = oracle.OracleWorld.view(small_world).frame(
        cam, wl.width, wl.height, 700.0, vg.VG_RAY_LOD, sun, 480.0, threads=THREADS)
    assert_same(gpu_hits(hits), ref, "small frame hits")
    occ = occ.numpy(np.uint8)
    assert np.array_equal(occ, ref_occ)
    assert occ[ref["hit"] == 0].sum() == 0
    assert 0 < occ.sum() < ref["hit"].sum()
    # the split entry points give the same frame
    h2 = small.trace_primary(cam, wl.width, wl.height, 700.0, vg.VG_RAY_LOD)
    o2 = small.trace_shadow(cam, wl.width, wl.height, h2, sun, 480.0)
    ctx.synchronize()
    assert_same(gpu_hits(h2), ref, "split primary")
    assert np.array_equal(o2.numpy(np.uint8), ref_occ)


@pytest.mark.parametrize("flags", [vg.VG_RAY_BEAM, vg.VG_RAY_LOD | vg.VG_RAY_BEAM])
def test_small_frame_with_beam_bit_exact(ctx, small, small_world, flags):
    # the fused beam prologue never changes a result: GPU beam frame == oracle beam frame ==
    # GPU frame without the beam; iteration counts (beam tile tests included) == oracle
    wl = workloads.SMALL
    _, cam = wl.camera()
    sun = workloads.LARGE.sun()
    print(f"case: small frame beam flags {flags}", flush=True)
    hits, occ, c = small.trace_frame_counts(cam, wl.width, wl.height, 700.0, flags, sun, 480.0)
    plain, plain_occ = small.trace_frame(cam, wl.width, wl.height, 700.0,
                                         flags & ~vg.VG_RAY_BEAM, sun, 480.0)
    ctx.synchronize()
    ref, ref_occ, rc = oracle.OracleWorld.view(small_world).frame(
        cam, wl.width, wl.height, 700.0, flags, sun, 480.0, threads=THREADS)
    assert_same(gpu_hits(hits), ref, "beam frame vs oracle")
    assert_same(gpu_hits(plain), ref, "beam frame vs plain frame")
    assert np.array_equal(occ.numpy(np.uint8), ref_occ)
    assert np.array_equal(plain_occ.numpy(np.uint8), ref_occ)
    assert c == rc and c["beam_test"] > 0


def mixed_rays(rng, n, dims):
    d = np.asarray(dims, np.float32)
    o = (rng.uniform(-0.2, 1.2, (n, 3)) * d).astype(np.float32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    k = rng.integers(0, 3, n)
    sel = rng.random(n) < 0.1
    dirs[sel, k[sel]] = 0.0
    sel = rng.random(n) < 0.02
    dirs[sel, k[sel]] = np.float32(5e-9)
    sel = rng.random(n) < 0.01
    dirs[sel] = 0.0
    tmax = rng.choice(np.array([32, 128, 512, 1024, np.inf], np.float32), n)
    flags = rng.choice(np.array([0, 1, 2, 4, 5], np.uint32), n)
    return oracle.make_rays(o, dirs, tmax, flags)


def test_batched_rays_bit_exact(ctx, small, small_world):
    rng = np.random.default_rng(11)
    rays = mixed_rays(rng, 1 << 18, small_world.dims)
    print(f"case: batched {len(rays)} mixed rays", flush=True)
    hits = small.trace_rays(ctx.rays(rays))
    ctx.synchronize()
    ref, _ = oracle.OracleWorld.view(small_world).trace_rays(rays, 0, THREADS)
    assert_same(gpu_hits(hits), ref, "batched")
    # a call-level flag ORs into every ray
    hits = small.trace_rays(ctx.rays(rays), flags=vg.VG_RAY_SKIP_TRANSPARENT)
    ctx.synchronize()
    ref, _ = oracle.OracleWorld.view(small_world).trace_rays(rays, oracle.SKIP_TRANSPARENT, THREADS)
    assert_same(gpu_hits(hits), ref, "batched skip-transparent")


def test_iteration_counts_match_oracle(ctx, small, small_world):
    wl = workloads.SMALL
    _, cam = wl.camera()
    sun = workloads.LARGE.sun()
    print("case: counts", flush=True)
    _, _, c = small.trace_frame_counts(cam, wl.width, wl.height, 700.0, vg.VG_RAY_LOD, sun, 480.0)
    _, _, rc = oracle.OracleWorld.view(small_world).frame(
        cam, wl.width, wl.height, 700.0, vg.VG_RAY_LOD, sun, 480.0, threads=THREADS)
    assert c == rc
    rng = np.random.default_rng(5)
    rays = mixed_rays(rng, 1 << 15, small_world.dims)
    _, c = small.trace_rays_counts(ctx.rays(rays))
    _, rc = oracle.OracleWorld.view(small_world).trace_rays(rays, 0, THREADS)
    assert c == rc


@pytest.mark.parametrize("origin", [(0, 0, 0), (96, 0, 160)])
def test_kat_floor_pillars(ctx, origin):
    ow = scenes.floor_pillars(origin)
    eyes, dirs = scenes.floor_pillar_rays(origin)
    rays = oracle.make_rays(eyes, dirs)
    print(f"case: floor+pillars origin {origin}", flush=True)
    dw = ctx.world(ow.dims).upload(scenes.to_world(ow))
    got = gpu_hits(dw.trace_rays(ctx.rays(rays)))
    ctx.synchronize()
    ref, _ = ow.trace_rays(rays)
    assert_same(got, ref, "pillars vs R1")
    r0 = np.array([ow.raycast_r0(e, d) for e, d in zip(eyes, dirs)])
    clean, floor, pillar, grazing, bad = scenes.classify(eyes, dirs, got, r0)
    assert not bad and clean > len(eyes) / 3 and floor > 0 and pillar > 0 and grazing < clean
    # the drop-in raycast returns PickHit with world voxel and normal
    from voxelg_amd.device import FACE_NORMALS, raycast
    for i in range(0, len(eyes), 37):
        p = raycast(eyes[i], dirs[i], dw, origin)
        if r0[i]["hit"]:
            assert p is not None and p.voxel == tuple(int(v) for v in ref[i]["voxel"])
            assert p.normal == FACE_NORMALS[int(ref[i]["face"])]
        else:
            assert p is None
    dw.close()


@pytest.mark.parametrize("origin", [(0, 0, 0), (96, 0, 160)])
def test_frame_on_pillars_folded_or_not(ctx, origin):
    # the frame kernel has one instantiation for windows whose origin folds an axis (toroidal
    # slots) and one for windows that fold none: both against R1's frame, bit for bit, with
    # LOD, the beam pre-pass and the sun shadow rays
    ow = scenes.floor_pillars(origin)
    cam = vg.Camera(pos=(origin[0] + 4.0, 110.0, origin[2] + 4.0), yaw=0.8, pitch=-0.6,
                    fov_y=np.radians(70.0)).uniform(320, 180)
    sun = workloads.LARGE.sun()
    flags = vg.VG_RAY_LOD | vg.VG_RAY_BEAM
    print(f"case: pillars frame origin {origin}", flush=True)
    dw = ctx.world(ow.dims).upload(scenes.to_world(ow))
    hits, occ = dw.trace_frame(cam, 320, 180, 700.0, flags, sun, 480.0)
    ctx.synchronize()
    got, occ = gpu_hits(hits), occ.numpy(np.uint8)
    ref, rsh, cnt = ow.frame(cam, 320, 180, 700.0, flags, sun, 480.0, threads=THREADS)
    assert_same(got, ref, f"pillars frame origin {origin} vs R1")
    assert occ.tobytes() == rsh.tobytes()
    assert ref["hit"].sum() > 10000 and 0 < rsh.sum() < ref["hit"].sum()
    assert cnt["beam_test"] > 0
    dw.close()


def test_kat_roof_shadow_and_empty(ctx):
    ow = scenes.roof_scene()
    origins, sun, under, open_ = scenes.roof_samples()
    rays = oracle.make_rays(origins, np.repeat(sun[None], len(origins), 0))
    print("case: roof shadow", flush=True)
    dw = ctx.world(ow.dims).upload(scenes.to_world(ow))
    got = gpu_hits(dw.trace_rays(ctx.rays(rays), flags=vg.VG_RAY_ANY_HIT))
    ctx.synchronize()
    ref, _ = ow.trace_rays(rays, oracle.ANY_HIT)
    assert_same(got, ref, "roof shadow")
    occ = got["hit"] == 1
    cpu = np.array([ow.raycast_r0(o, sun)["hit"] for o in origins]) == 1
    assert np.sum((occ != cpu) & (under | open_)) == 0
    assert np.sum(under & occ) > 20 and np.sum(open_ & ~occ) > 20
    dw.close()
    empty = vg.World(scenes.REF_DIMS)
    dw = ctx.world(empty.dims).upload(empty)
    got = gpu_hits(dw.trace_rays(ctx.rays(rays)))
    ctx.synchronize()
    assert not got["hit"].any() and np.all(got["t"] == -1.0) and np.all(got["face"] == 6)
    dw.close()


@pytest.mark.parametrize("schedule", ["static", (8, 1), (16, 32), (32, 64), (4, 16),
                                      (16, 4, 1), (12, 8, 7), (20, 16, 64), (16, 4, 4, 4),
                                      (16, 8, 2, 16), (12, 4, 8, 32)])
def test_schedules_agree(ctx, small, small_world, schedule):
    # the one-thread-per-ray kernels and the persistent ray-compacting kernels at several
    # (waves per CU, refill threshold) settings return the same bytes as the oracle
    from voxelg_amd import _lib as L
    if schedule == "static":
        ctx.set_option(L.VG_OPT_SCHEDULE, L.VG_SCHEDULE_STATIC)
    else:
        ctx.set_option(L.VG_OPT_SCHEDULE, L.VG_SCHEDULE_PERSISTENT)
        for o, v in ((L.VG_OPT_FRAME_WAVES_PER_CU, schedule[0]), (L.VG_OPT_FRAME_REFILL, schedule[1]),
                     (L.VG_OPT_RAYS_WAVES_PER_CU, schedule[0]), (L.VG_OPT_RAYS_REFILL, schedule[1])):
            ctx.set_option(o, v)
        # work stripes of the frame kernel (1: one counter; 64: more stripes than the small
        # frame's rows of blocks hold evenly)
        ctx.set_option(L.VG_OPT_FRAME_STRIPES, schedule[2] if len(schedule) > 2 else 4)
        # pixel-block shape (width x 64 / width): the beam marches one box per block
        ctx.set_option(L.VG_OPT_FRAME_BLOCK_W, schedule[3] if len(schedule) > 3 else 8)
    try:
        wl = workloads.SMALL
        _, cam = wl.camera()
        sun = workloads.LARGE.sun()
        print(f"case: schedule {schedule}", flush=True)
        flags = vg.VG_RAY_LOD | vg.VG_RAY_BEAM
        hits, occ, c = small.trace_frame_counts(cam, wl.width, wl.height, 700.0, flags, sun, 480.0)
        ref, ref_occ, rc = oracle.OracleWorld.view(small_world).frame(
            cam, wl.width, wl.height, 700.0, flags, sun, 480.0, threads=THREADS)
        assert_same(gpu_hits(hits), ref, f"frame {schedule}")
        assert np.array_equal(occ.numpy(np.uint8), ref_occ)
        if ctx.get_option(L.VG_OPT_FRAME_BLOCK_W) == 8 or schedule == "static":
            assert c == rc  # the oracle's beam marches 8x8 blocks
        rays = mixed_rays(np.random.default_rng(9), 100_000, small_world.dims)
        got, c = small.trace_rays_counts(ctx.rays(rays))
        ref, rc = oracle.OracleWorld.view(small_world).trace_rays(rays, 0, THREADS)
        assert_same(gpu_hits(got), ref, f"rays {schedule}")
        assert c == rc
    finally:
        ctx.set_option(L.VG_OPT_SCHEDULE, L.VG_SCHEDULE_PERSISTENT)
        for o in (L.VG_OPT_FRAME_WAVES_PER_CU, L.VG_OPT_RAYS_WAVES_PER_CU):
            ctx.set_option(o, 16)
        for o in (L.VG_OPT_FRAME_REFILL, L.VG_OPT_RAYS_REFILL):
            ctx.set_option(o, 32)
        ctx.set_option(L.VG_OPT_FRAME_STRIPES, 4)
        ctx.set_option(L.VG_OPT_FRAME_BLOCK_W, 8)


def test_autotune_keeps_results(ctx, small, small_world):
    from voxelg_amd import _lib as L
    wl = workloads.SMALL
    _, cam = wl.camera()
    sun = workloads.LARGE.sun()
    flags = vg.VG_RAY_LOD | vg.VG_RAY_BEAM
    opts = (L.VG_OPT_FRAME_WAVES_PER_CU, L.VG_OPT_FRAME_REFILL, L.VG_OPT_FRAME_STRIPES,
            L.VG_OPT_FRAME_BLOCK_W)
    before = [ctx.get_option(o) for o in opts]
    try:
        ms = small.autotune_frame(cam, wl.width, wl.height, 700.0, flags, sun, 480.0)
        assert 0.0 < ms < 100.0
        assert 1 <= ctx.get_option(L.VG_OPT_FRAME_WAVES_PER_CU) <= 64
        assert 1 <= ctx.get_option(L.VG_OPT_FRAME_REFILL) <= 64
        assert 1 <= ctx.get_option(L.VG_OPT_FRAME_STRIPES) <= 64
        assert ctx.get_option(L.VG_OPT_FRAME_BLOCK_W) in (4, 8, 16, 32)
        assert ctx.get_option(L.VG_OPT_CU_COUNT) >= 1
        hits, occ = small.trace_frame(cam, wl.width, wl.height, 700.0, flags, sun, 480.0)
        ref, ref_occ, _ = oracle.OracleWorld.view(small_world).frame(
            cam, wl.width, wl.height, 700.0, flags, sun, 480.0, threads=THREADS)
        ctx.synchronize()
        assert_same(gpu_hits(hits), ref, "frame after autotune")
        assert np.array_equal(occ.numpy(np.uint8), ref_occ)
        with pytest.raises(Exception):
            ctx.set_option(L.VG_OPT_FRAME_REFILL, 0)
        with pytest.raises(Exception):
            ctx.set_option(L.VG_OPT_FRAME_BLOCK_W, 12)
    finally:
        for o, v in zip(opts, before):
            ctx.set_option(o, v)


def batched_setup(ctx, small, small_world):
    from voxelg_amd.device import DevicePtr
    rays, sizes = workloads.BATCHED.rays(small_world)
    rbuf = ctx.rays(rays)
    hbuf = ctx.alloc(len(rays) * 20)
    batches, off = [], 0
    for n in sizes:
        batches.append((DevicePtr(rbuf, off * 32), DevicePtr(hbuf, off * 20), n, 0))
        off += n
    return rays, hbuf, small.batch_set(batches)


def test_uneven_batches_bit_exact(ctx, small, small_world):
    rays, hbuf, bs = batched_setup(ctx, small, small_world)
    print(f"case: {bs.n} uneven batches, {bs.total} rays", flush=True)
    assert bs.n == 96 and bs.total == len(rays)
    small.trace_batches(bs)
    ctx.synchronize()
    ref, rc = oracle.OracleWorld.view(small_world).trace_rays(rays, 0, THREADS)
    assert_same(gpu_hits(hbuf), ref, "uneven batches")
    assert small.trace_batches_counts(bs) == rc
    assert_same(gpu_hits(hbuf), ref, "uneven batches (counting pass)")


def test_graph_replays_match_eager(ctx, small, small_world):
    from voxelg_amd.device import Graph
    rays, hbuf, bs = batched_setup(ctx, small, small_world)
    stream = ctx.stream()
    small.trace_batches(bs, stream)
    stream.synchronize()
    eager = gpu_hits(hbuf).copy()
    g = Graph(ctx, stream, lambda s: small.trace_batches(bs, s))
    other = mixed_rays(np.random.default_rng(4), 50_000, small_world.dims)
    for rep in range(3):
        print(f"case: graph replay {rep}", flush=True)
        hbuf.zero()
        small.trace_rays(ctx.rays(other))  # eager work between replays
        ctx.synchronize()
        g.launch(stream)
        stream.synchronize()
        assert gpu_hits(hbuf).tobytes() == eager.tobytes(), f"replay {rep}"


def test_edge_cases_and_errors(ctx, small, small_world):
    from voxelg_amd._lib import VGError
    # zero rays is a no-op
    small.trace_rays(ctx.alloc(32), n=0)
    # a world that was never uploaded refuses to trace
    dw = ctx.world((64, 64, 64))
    with pytest.raises(VGError):
        dw.trace_rays(ctx.alloc(4 * 32), n=4)
    # a folding origin needs dims that are multiples of 256
    w = vg.World((64, 64, 64), origin=(32, 0, 0))
    with pytest.raises(VGError):
        dw.upload(w)
    dw.close()
    # camera resolution must match the launch
    _, cam = workloads.SMALL.camera()
    with pytest.raises(VGError):
        small.trace_primary(cam, 640, 360, 512.0)


def test_small_frame_vs_r0_full(ctx, small, small_world):
    # the R0 agreement criterion on the 1280x720 frame of the 256x128x256 world (primary tmax
    # 512, LOD off) plus its sun shadow rays: the GPU's hits against the reference CPU
    # traversal R0 (src/raycast.rs:21-141) traced over the same rays, every pixel
    wl = workloads.SMALL
    _, cam = wl.camera()
    sun = workloads.LARGE.sun()
    print("case: small frame vs R0", flush=True)
    hits, occ = small.trace_frame(cam, wl.width, wl.height, wl.primary_tmax, 0, sun, 480.0)
    ctx.synchronize()
    got, occ = gpu_hits(hits), occ.numpy(np.uint8)
    ow = oracle.OracleWorld.view(small_world)
    r0, s0 = ow.frame_r0(cam, wl.width, wl.height, wl.primary_tmax, sun, 480.0, threads=THREADS)
    dirs = oracle.primary_dirs(cam, wl.width, wl.height)
    c = oracle.classify_vs_r0(cam[:3], dirs, got, r0, tmax=wl.primary_tmax)
    scenes.assert_agrees_with_r0(c, "small frame GPU vs R0")
    assert c["agree_clean"] > 0.85 * r0["hit"].sum()
    alln = np.ones(len(got), bool)
    cs = oracle.classify_shadows_vs_r0(ow, cam[:3], dirs, got, occ, s0, sun, 480.0,
                                       scenes.same_shadow_ray(got, r0, cam[:3], dirs, alln))
    print("small frame shadows GPU vs R0", cs, flush=True)
    assert cs["clean_disagree"] == 0 and cs["rays"] > 0.8 * c["agree"], cs


def test_uneven_batches_vs_r0_full(ctx, small, small_world):
    # one whole step of the 96 uneven batches against R0 (which normalizes each direction
    # itself, as vg_trace_batches does)
    rays, hbuf, bs = batched_setup(ctx, small, small_world)
    print(f"case: {bs.n} uneven batches vs R0", flush=True)
    small.trace_batches(bs)
    ctx.synchronize()
    got = gpu_hits(hbuf)
    r0 = oracle.OracleWorld.view(small_world).raycast_rays(rays, THREADS)
    c = oracle.classify_vs_r0(rays["o"], oracle.normalize_or_zero_v(rays["d"]), got, r0,
                              tmax=rays["tmax"])
    scenes.assert_agrees_with_r0(c, "uneven batches GPU vs R0")
    assert c["agree_clean"] > 0.8 * r0["hit"].sum()


def test_large_frame_bit_exact_and_vs_r0(ctx):
    # the headline frame (2048x256x2048 world, 1920x1080, LOD 400 / 520, beam, sun shadow per
    # hit): every pixel equals the oracle's R1 bit for bit, and against R0 the hits before
    # t 400 (where LOD cannot have fired) and their shadow rays meet the R0 agreement criterion
    wl = workloads.LARGE
    print("case: large world generate", flush=True)
    world = wl.world()
    _, cam = wl.camera()
    sun = wl.sun()
    dw = ctx.world(world.dims).upload(world)
    print("case: large frame", flush=True)
    hits, occ = dw.trace_frame(cam, wl.width, wl.height, wl.primary_tmax, wl.primary_flags, sun,
                               wl.shadow_tmax)
    ctx.synchronize()
    flat = gpu_hits(hits)
    occ = occ.numpy(np.uint8)
    ow = oracle.OracleWorld.view(world)
    print("case: large frame oracle R1", flush=True)
    ref, ref_occ, _ = ow.frame(cam, wl.width, wl.height, wl.primary_tmax, wl.primary_flags,
                               sun, wl.shadow_tmax, threads=THREADS)
    assert_same(flat, ref, "large frame")
    assert np.array_equal(occ, ref_occ)
    # the beam prologue is a pure acceleration: the whole frame without it is identical
    assert wl.primary_flags & vg.VG_RAY_BEAM
    h2, o2 = dw.trace_frame(cam, wl.width, wl.height, wl.primary_tmax,
                            wl.primary_flags & ~vg.VG_RAY_BEAM, sun, wl.shadow_tmax)
    ctx.synchronize()
    assert gpu_hits(h2).tobytes() == flat.tobytes()
    assert np.array_equal(o2.numpy(np.uint8), occ)
    h = flat["hit"] == 1
    assert h.mean() > 0.3
    assert np.all(flat["t"][h] <= wl.primary_tmax) and np.all(flat["t"][h] >= 0)
    assert np.all(occ[~h] == 0)
    dw.close()
    print("case: large frame R0", flush=True)
    r0, s0 = ow.frame_r0(cam, wl.width, wl.height, wl.primary_tmax, sun, wl.shadow_tmax,
                         threads=THREADS)
    dirs = oracle.primary_dirs(cam, wl.width, wl.height)
    sel = (flat["hit"] == 0) | (flat["t"] < 400.0)
    c = oracle.classify_vs_r0(cam[:3], dirs[sel], flat[sel], r0[sel], tmax=wl.primary_tmax)
    scenes.assert_agrees_with_r0(c, "large frame GPU vs R0 (t < 400)")
    assert c["agree_clean"] > 0.5 * sel.sum()
    cs = oracle.classify_shadows_vs_r0(ow, cam[:3], dirs, flat, occ, s0, sun, wl.shadow_tmax,
                                       scenes.same_shadow_ray(flat, r0, cam[:3], dirs, sel))
    print("large frame shadows GPU vs R0", cs, flush=True)
    assert cs["clean_disagree"] == 0 and cs["rays"] > 0.8 * c["agree"], cs
