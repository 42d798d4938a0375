"""Benchmark: Mrays/s of one 1080p primary + sun-shadow frame on the 2048x256x2048 world.

One step = one frame: 1920x1080 pinhole primary rays (tmax 700, brick LOD past 400 and tile
LOD past 520) through the seeded benchmark world, then one any-hit shadow ray toward the sun
(tmax 480) from every primary hit, all in one launch of the engine's frame kernel. The world
is generated on the host, uploaded once, and resident in HBM before the timed region.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
  python bench.py --workload batched   # uneven batched queries (96 batches per step)
  python bench.py --workload gi        # GI probe update: plain 256-ray vs learned 16-ray
  python bench.py --workload dynamic   # 4096 sphere edits per frame + retrace + probe update

Prints ONE JSON line on stdout; details go to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
# value: the frame's Mrays/s; the "gi" object holds the probe-update half (ms per update)
METRIC = "Mrays/s at 1080p primary+shadow; GI probe ms at \u22642% irradiance err vs plain"
GI_ERR_LIMIT = 0.02


def gi_threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(world, wl, cam, sun, gpu_hits, gpu_occ, row_budget_s=2.0):
    """The reference's CPU traversal on this host, after benches/trace.rs:24-41 (warm-up, then
    repeated timed runs): R0 (src/raycast.rs:21-141 restated, the per-voxel DDA) and R1 (the
    exact hierarchical walk the GPU reproduces), each from a -O3 -march=native build of the
    oracle made on this host, on every host thread this process may use (at most 16, one
    GPU's share), rows interleaved over the threads. Timed on every k-th row of the frame (k
    sized so one R0 run takes about row_budget_s): one warm-up, then the median of 5 runs.
    Before the timing, R0 and R1 trace the whole frame once: the GPU frame is compared with
    both (the parity report)."""
    import numpy as np

    import oracle
    L, build = oracle.timing_lib()
    cpu = oracle.host_cpu()
    threads = max(1, min(16, cpu["usable_cpus"]))
    ow = oracle.OracleWorld.view(world)
    W, H = wl.width, wl.height
    args = (cam, W, H, wl.primary_tmax)
    t0 = time.perf_counter()
    r0, s0 = ow.frame_r0(*args, sun, wl.shadow_tmax, threads=threads, L=L)
    full_r0_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    r1, s1, _ = ow.frame(*args, wl.primary_flags, sun, wl.shadow_tmax, threads=threads, L=L)
    full_r1_s = time.perf_counter() - t0
    parity = parity_report(wl, cam, sun, ow, gpu_hits, gpu_occ, r0, s0, r1, s1)

    step = max(1, int(np.ceil(full_r0_s / row_budget_s)))
    rows = (step // 2, H, step)
    n_rows = len(range(*rows))

    def timed(fn):
        fn()  # warm-up
        ts, rays = [], 0
        for _ in range(5):
            t0 = time.perf_counter()
            h = fn()
            ts.append(time.perf_counter() - t0)
            rays = n_rows * W + int(h["hit"].sum())
        return rays, float(np.median(ts)), ts

    n0, med0, ts0 = timed(lambda: ow.frame_r0(*args, sun, wl.shadow_tmax, rows, threads,
                                             L=L)[0])
    n1, med1, ts1 = timed(lambda: ow.frame(*args, wl.primary_flags, sun, wl.shadow_tmax, rows,
                                          threads, L=L)[0])
    return {"value": n0 / med0 / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{n_rows} of {H} rows (every {step}th), {n0} primary+shadow rays per run; "
                      f"1 warm-up + median of 5 runs ({', '.join(f'{t:.2f}' for t in ts0)} s); "
                      f"R0 = src/raycast.rs per-voxel DDA, one raycast() per ray",
            "r1_value": n1 / med1 / 1e6,
            "r1_sample": f"same rows, median of 5 runs ({', '.join(f'{t:.3f}' for t in ts1)} s); "
                         f"R1 = the exact hierarchical walk (LOD 400/520, as the GPU frame)",
            "build": build, "threads": threads, "cpu_model": cpu["model"],
            "physical_cores": cpu["physical_cores"], "logical_cpus": cpu["logical_cpus"],
            "usable_cpus": cpu["usable_cpus"],
            "full_frame_s": {"r0": round(full_r0_s, 2), "r1": round(full_r1_s, 2)}}, parity


def parity_report(wl, cam, sun, ow, got, occ, r0, s0, r1, s1):
    """The GPU frame against the CPU traversals over every pixel: bit for bit against R1 (the
    exact walk with the same LOD), and against R0 (src/raycast.rs) by the rule of the
    reference's accel_matches_cpu_raycast test (src/accel.rs:655-718) where the LOD frame
    equals the exact walk (rays without a hit at t >= 400), plus the shadow rays of the hits
    that agree on voxel, face and t."""
    import numpy as np

    import oracle
    dirs = oracle.primary_dirs(cam, wl.width, wl.height)
    occ = np.asarray(occ).reshape(-1)
    r1_same = got.tobytes() == r1.tobytes() and np.array_equal(occ, s1)
    sel = (got["hit"] == 0) | (got["t"] < 400.0)
    c = oracle.classify_vs_r0(cam[:3], dirs[sel], got[sel], r0[sel], tmax=wl.primary_tmax)
    same = sel & (got["hit"] != 0) & (r0["hit"] != 0) & \
        np.all(got["voxel"] == r0["voxel"], axis=1) & (got["face"] == r0["face"]) & \
        (got["t"] == r0["t"])
    cs = oracle.classify_shadows_vs_r0(ow, cam[:3], dirs, got, occ, s0, sun, wl.shadow_tmax,
                                       same)
    return {"vs_r1_bit_exact": bool(r1_same), "rays": int(len(got)),
            "vs_r0": {"compared": c["rays"], "outside_lod_region": int((~sel).sum()),
                      "agree": c["agree"], "agree_clean": c["agree_clean"],
                      "ties": c["ties"], "edge_tie_disagreements": c["disagree_ties"],
                      "tmax_ties": c["tmax_ties"], "clean_disagree": c["clean_disagree"],
                      "face_mismatch": c["face_mismatch"],
                      "material_mismatch": c["material_mismatch"], "max_dt": c["max_dt"],
                      "shadow": cs}}


def parity_configs(ctx, threads):
    """The other two traversal workloads against the CPU traversals, outside the timed
    region: the small frame (1280x720 over the 256x128x256 world, primary tmax 512, LOD off,
    plus a sun shadow ray per hit, tmax 480) and the batched step (the 96 uneven batches
    over the same world): the GPU's records bit for bit against R1, and against R0
    (src/raycast.rs) by the rule of the reference's accel_matches_cpu_raycast test
    (src/accel.rs:655-718), with the edge-tie and tmax-tie counts."""
    import numpy as np

    import oracle
    from voxelg_amd import workloads
    from voxelg_amd.device import DevicePtr, decode_hits

    def summary(c):
        return {"compared": c["rays"], "agree": c["agree"], "agree_clean": c["agree_clean"],
                "ties": c["ties"], "edge_tie_disagreements": c["disagree_ties"],
                "tmax_ties": c["tmax_ties"], "clean_disagree": c["clean_disagree"],
                "face_mismatch": c["face_mismatch"], "material_mismatch": c["material_mismatch"],
                "max_dt": c["max_dt"]}

    wl = workloads.SMALL
    world = wl.world(threads)
    dw = ctx.world(world.dims).upload(world)
    ow = oracle.OracleWorld.view(world)
    _, cam = wl.camera()
    sun, stmax = workloads.LARGE.sun(), 480.0
    W, H = wl.width, wl.height
    hits, occ = dw.trace_frame(cam, W, H, wl.primary_tmax, 0, sun, stmax)
    ctx.synchronize()
    got, occ = decode_hits(hits), occ.numpy(np.uint8)
    r1, s1, _ = ow.frame(cam, W, H, wl.primary_tmax, 0, sun, stmax, threads=threads)
    r0, s0 = ow.frame_r0(cam, W, H, wl.primary_tmax, sun, stmax, threads=threads)
    dirs = oracle.primary_dirs(cam, W, H)
    c1 = oracle.classify_vs_r0(cam[:3], dirs, got, r0, tmax=wl.primary_tmax)
    same = (got["hit"] != 0) & (r0["hit"] != 0) & np.all(got["voxel"] == r0["voxel"], axis=1) \
        & (got["face"] == r0["face"]) & (got["t"] == r0["t"])
    cs = oracle.classify_shadows_vs_r0(ow, cam[:3], dirs, got, occ, s0, sun, stmax, same)
    out = {"small_frame_1280x720": {"rays": int(len(got)),
                       "vs_r1_bit_exact": bool(got.tobytes() == r1.tobytes() and
                                               np.array_equal(occ, s1)),
                       "vs_r0": dict(summary(c1), shadow=cs)}}
    rays, sizes = workloads.BATCHED.rays(world)
    rbuf = ctx.rays(rays)
    hbuf = ctx.alloc(len(rays) * 20)
    batches, off = [], 0
    for n in sizes:
        batches.append((DevicePtr(rbuf, off * 32), DevicePtr(hbuf, off * 20), n, 0))
        off += n
    bs = dw.batch_set(batches)
    dw.trace_batches(bs)
    ctx.synchronize()
    got3 = decode_hits(hbuf)
    ref3, _ = ow.trace_rays(rays, 0, threads)
    r03 = ow.raycast_rays(rays, threads)
    c3 = oracle.classify_vs_r0(rays["o"], oracle.normalize_or_zero_v(rays["d"]), got3, r03,
                               tmax=rays["tmax"])
    out["batched_96"] = {"rays": int(len(rays)), "batches": int(bs.n),
                      "vs_r1_bit_exact": bool(got3.tobytes() == ref3.tobytes()),
                      "vs_r0": summary(c3)}
    dw.close()
    return out


TRAFFIC_PROFILE = "profiles/k_frame_persistent_large_1080p_pmc.md"
BATCHED_TRAFFIC_PROFILE = "profiles/k_batches_persistent_batched_pmc.md"
GI_TRAFFIC_PROFILE = "profiles/k_probe_paths_learned_gi_pmc.md"


COUNTERS_PROFILE = "profiles/cache_hit_rates_and_mfma_util_pmc.md"


def profile_value(profile, key):
    """A figure recorded as `key: value` in a committed rocprofv3 --pmc summary (counters
    cannot be read from inside the timed process itself); None when absent."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), profile)
    try:
        m = re.search(r"^" + re.escape(key) + r": ([0-9.eE+-]+)$", open(path).read(), re.M)
    except OSError:
        return None
    return float(m.group(1)) if m else None


def measured_traffic(profile=None, key="traffic_bytes_per_frame"):
    """DRAM-side bytes per launch of a workload's dominant kernel, from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs (copy-calibrated); None when the summary is
    absent. Default: k_beam + k_frame_persistent per headline frame."""
    v = profile_value(profile or TRAFFIC_PROFILE, key)
    return None if v is None else int(v)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["frame", "batched", "gi", "dynamic"], default="frame")
    ap.add_argument("--check-every", type=int, default=16,
                    help="dynamic: bit-exact oracle check of the frame every N frames")
    ap.add_argument("--gi-train-updates", type=int, default=150,
                    help="learned GI: online-training updates before the timed ones")
    ap.add_argument("--gi-steps", type=int, default=20, help="GI: timed probe updates per mode")
    ap.add_argument("--no-gi", action="store_true", help="frame: skip the GI probe half")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (a plain copy kernel "
                         "of known bytes, to scale rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the ray-query path runs on a single MI355X "
              "(rays are independent: frames would shard by screen blocks, probe updates by "
              "probe slabs).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "gi":
        return bench_gi_main(args)
    if args.workload == "dynamic":
        return bench_dynamic(args)

    import numpy as np

    from voxelg_amd import workloads
    from voxelg_amd._lib import HIT_RECORD_BYTES, MATERIAL_BYTES, algorithmic_bytes
    from voxelg_amd.device import Context, decode_hits, device_count

    assert device_count() > 0, "bench.py needs a HIP device"
    wl = workloads.LARGE
    t0 = time.perf_counter()
    world = wl.world()
    gen_s = time.perf_counter() - t0
    log(f"world {wl.dims} generated in {gen_s:.1f} s, solid {world.solid_fraction():.3f}")
    _, cam = wl.camera()
    sun = wl.sun()
    ctx = Context(0)
    t0 = time.perf_counter()
    dw = ctx.world(wl.dims).upload(world)
    log(f"uploaded in {time.perf_counter() - t0:.2f} s")

    W, H = wl.width, wl.height
    hits = ctx.alloc(W * H * 20)
    occ = ctx.alloc(W * H)
    stream = ctx.stream()

    def frame():
        dw.trace_frame(cam, W, H, wl.primary_tmax, wl.primary_flags, sun, wl.shadow_tmax, hits,
                       occ, stream)

    # schedule autotuning (outside the timed region): pixel-block shape x waves per CU x refill
    # threshold x work stripes of the persistent frame kernel, timed on this very frame
    from voxelg_amd import _lib as L
    tuned_ms = dw.autotune_frame(cam, W, H, wl.primary_tmax, wl.primary_flags, sun,
                                 wl.shadow_tmax)
    sched = {"waves_per_cu": ctx.get_option(L.VG_OPT_FRAME_WAVES_PER_CU),
             "refill_below": ctx.get_option(L.VG_OPT_FRAME_REFILL),
             "stripes": ctx.get_option(L.VG_OPT_FRAME_STRIPES),
             "block": [ctx.get_option(L.VG_OPT_FRAME_BLOCK_W),
                       64 // ctx.get_option(L.VG_OPT_FRAME_BLOCK_W)],
             "cus": ctx.get_option(L.VG_OPT_CU_COUNT)}
    log(f"autotuned persistent frame kernel: {sched} ({tuned_ms:.3f} ms)")

    # algorithmic accounting (outside the timed region): iteration classes of this frame
    _, _, counts = dw.trace_frame_counts(cam, W, H, wl.primary_tmax, wl.primary_flags, sun,
                                         wl.shadow_tmax)
    frame()
    stream.synchronize()
    hh = decode_hits(hits)
    n_hit = int(hh["hit"].sum())
    n_shadowed = int(occ.numpy(np.uint8).sum())
    rays_per_frame = W * H + n_hit
    bytes_per_frame = (algorithmic_bytes(counts) + W * H * HIT_RECORD_BYTES +
                       n_hit * MATERIAL_BYTES + W * H)
    log(f"frame: {W}x{H} primary, {n_hit} hits -> {n_hit} shadow rays ({n_shadowed} occluded)")
    log(f"iteration classes: {counts}")

    for _ in range(args.warmup):
        frame()
    stream.synchronize()
    starts = [ctx.event() for _ in range(args.steps)]
    ends = [ctx.event() for _ in range(args.steps)]
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        frame()
        ends[i].record(stream)
    stream.synchronize()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [s.elapsed_ms(e) for s, e in zip(starts, ends)]
    avg_kern_s = float(np.mean(kern_ms)) / 1e3
    ms_per_step = wall / args.steps * 1e3
    value = rays_per_frame * args.steps / wall / 1e6
    achieved = bytes_per_frame / avg_kern_s / 1e9
    log(f"wall {ms_per_step:.3f} ms/frame, kernel {np.mean(kern_ms):.3f} ms "
        f"(min {np.min(kern_ms):.3f}), {value:.0f} Mrays/s")

    if args.copy_calibration_mib:
        import ctypes
        from voxelg_amd._lib import check, lib
        nbytes = args.copy_calibration_mib << 20
        a, b = ctx.alloc(nbytes).zero(), ctx.alloc(nbytes)
        check(lib().vg_copy_device(ctx._h, b.data_ptr(), a.data_ptr(), ctypes.c_size_t(nbytes)),
              "vg_copy_device")
        ctx.synchronize()
        log(f"copy calibration: {nbytes} bytes device-to-device")

    # the water secondary rays of the same frame (reflection + refraction of every water
    # pixel, SKIP_TRANSPARENT, tmax 200), outside the timed region
    water = (hh["hit"] == 1) & (hh["material"] >= 5) & (hh["material"] <= 12)
    refl, refr = dw.trace_water_secondary(cam, W, H, hits, stream=stream)
    wa, wb = ctx.event(), ctx.event()
    wa.record(stream)
    for _ in range(10):
        dw.trace_water_secondary(cam, W, H, hits, refl, refr, stream)
    wb.record(stream)
    stream.synchronize()
    water_ms = wa.elapsed_ms(wb) / 10
    water_out = {"pixels": int(water.sum()), "rays": 2 * int(water.sum()),
                 "ms": round(water_ms, 4),
                 "mrays_per_s": round(2 * int(water.sum()) / water_ms / 1e3, 1)}
    log(f"water secondary rays: {water_out}")

    # the same world generated on the device (no host arrays): its time, and the frame it
    # gives against the uploaded host world's, outside the timed region
    gw = ctx.world(wl.dims)
    ga, gb = ctx.event(), ctx.event()
    ctx.synchronize()
    t0 = time.perf_counter()
    ga.record(stream)
    gw.generate(wl.seed, (0, 0, 0), stream)
    gb.record(stream)
    stream.synchronize()
    gen_wall = time.perf_counter() - t0
    ghits, gocc = gw.trace_frame(cam, W, H, wl.primary_tmax, wl.primary_flags, sun,
                                 wl.shadow_tmax)
    ctx.synchronize()
    gen_out = {"ms": round(ga.elapsed_ms(gb), 3), "wall_ms": round(gen_wall * 1e3, 3),
               "host_generate_s": round(gen_s, 3),
               "frame_identical": bool(decode_hits(ghits).tobytes() == hh.tobytes() and
                                       gocc.numpy(np.uint8).tobytes() ==
                                       occ.numpy(np.uint8).tobytes())}
    gw.close()
    log(f"device world generation: {gen_out}")

    # the metric's second half: the GI probe update on the same world, learned 16-ray (cache
    # trained online) against plain 256-ray probes, errors against the FP64 reference
    gi_out = None
    if not args.no_gi:
        gi_out = gi_probe_bench(ctx, dw, world, sun, args.gi_steps, args.gi_train_updates,
                                gi_threads())

    base = parity = None
    if not args.no_cpu_baseline:
        base, parity = cpu_baseline(world, wl, cam, sun, hh, occ.numpy(np.uint8))
        log(f"cpu baseline: {base}")
        parity["other_configs"] = parity_configs(ctx, base["threads"])
        log(f"parity: {parity}")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm terrain with caves, generated on the host)",
        "config": {
            "workload": "large: 2048x256x2048 world (32x4x32 chunks), 1920x1080 primary rays "
                        "tmax 700 with brick/tile LOD past 400/520, one sun shadow ray per hit "
                        "tmax 480",
            "world": list(wl.dims), "resolution": [W, H], "seed": wl.seed,
            "primary_rays": W * H, "shadow_rays": n_hit,
            "schedule": dict(sched, kernel="persistent ray-compacting + beam pre-pass"),
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
            "traffic_source": TRAFFIC_PROFILE if measured_traffic() is not None else None,
            "l2_hit_rate": profile_value(COUNTERS_PROFILE, "l2_hit_rate_frame"),
            "l2_hit_rate_source": COUNTERS_PROFILE,
            "kernel": "k_beam + k_frame_persistent", "kernel_ms": round(avg_kern_s * 1e3, 4),
            "algorithmic_bytes": int(bytes_per_frame),
        },
        "cpu_baseline": base,
        "parity": parity,
        "gi": gi_out,
        "water_secondary": water_out,
        "device_world_generation": gen_out,
    }
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """One step = 96 uneven batches (sizes log-uniform in [256, 262144]) of mixed primary,
    shadow and AO rays on the 256x128x256 world, replayed as one captured HIP graph."""
    import numpy as np

    import voxelg_amd as vg
    from voxelg_amd import _lib as L
    from voxelg_amd import workloads
    from voxelg_amd._lib import (HIT_RECORD_BYTES, MATERIAL_BYTES, RAY_INPUT_BYTES,
                                 algorithmic_bytes)
    from voxelg_amd.device import Context, DevicePtr, Graph, decode_hits, device_count

    assert device_count() > 0, "bench.py needs a HIP device"
    wl = workloads.BATCHED
    world = wl.world()
    rays, sizes = wl.rays(world)
    ctx = Context(0)
    dw = ctx.world(world.dims).upload(world)
    rbuf = ctx.rays(rays)
    hbuf = ctx.alloc(len(rays) * 20)
    batches, off = [], 0
    for n in sizes:
        batches.append((DevicePtr(rbuf, off * 32), DevicePtr(hbuf, off * 20), n, 0))
        off += n
    bs = dw.batch_set(batches)
    tuned = dw.autotune_batches(bs)
    sched = {"waves_per_cu": ctx.get_option(L.VG_OPT_RAYS_WAVES_PER_CU),
             "refill_below": ctx.get_option(L.VG_OPT_RAYS_REFILL),
             "cus": ctx.get_option(L.VG_OPT_CU_COUNT)}
    log(f"{bs.n} batches, {bs.total} rays; autotuned {sched} ({tuned:.3f} ms)")
    counts = dw.trace_batches_counts(bs)
    hits = decode_hits(hbuf)
    n_hit = int(hits["hit"].sum())
    bytes_per_step = (algorithmic_bytes(counts) + bs.total * (RAY_INPUT_BYTES + HIT_RECORD_BYTES)
                      + n_hit * MATERIAL_BYTES)
    stream = ctx.stream()
    graph = Graph(ctx, stream, lambda st: dw.trace_batches(bs, st))
    for _ in range(args.warmup):
        graph.launch(stream)
    stream.synchronize()
    ev = [(ctx.event(), ctx.event()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        graph.launch(stream)
        b.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [a.elapsed_ms(b) for a, b in ev]
    value = bs.total * args.steps / wall / 1e6
    achieved = bytes_per_step / (float(np.mean(kern_ms)) / 1e3) / 1e9
    log(f"step {wall / args.steps * 1e3:.3f} ms wall, graph {np.mean(kern_ms):.3f} ms, "
        f"{value:.0f} Mrays/s, {n_hit} hits")
    base = None
    if not args.no_cpu_baseline:
        # R0 (one raycast() per ray) and R1 over whole steps, the host-native -O3 build, every
        # usable host thread (at most 16): one warm-up, then the median of 5 steps
        import oracle
        L, build = oracle.timing_lib()
        cpu = oracle.host_cpu()
        threads = gi_threads()
        ow = oracle.OracleWorld.view(world)

        def med5(fn):
            fn()
            ts = []
            for _ in range(5):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return float(np.median(ts)), ts
        m0, ts0 = med5(lambda: ow.raycast_rays(rays, threads, L=L))
        m1, ts1 = med5(lambda: ow.trace_rays(rays, 0, threads, L=L))
        r0 = ow.raycast_rays(rays, threads, L=L)
        c = oracle.classify_vs_r0(rays["o"], oracle.normalize_or_zero_v(rays["d"]), hits, r0,
                                  tmax=rays["tmax"])
        base = {"value": len(rays) / m0 / 1e6, "unit": "Mrays/s", "cores": threads,
                "kind": "port",
                "sample": f"whole steps ({len(rays)} rays), 1 warm-up + median of 5 "
                          f"({', '.join(f'{t:.2f}' for t in ts0)} s), R0 = src/raycast.rs "
                          f"per-voxel DDA",
                "r1_value": len(rays) / m1 / 1e6, "build": build, "cpu_model": cpu["model"],
                "physical_cores": cpu["physical_cores"], "usable_cpus": cpu["usable_cpus"]}
        parity = {"vs_r0": {k: v for k, v in c.items() if k != "bad_index"}}
        log(f"cpu baseline: {base}")
        log(f"parity: {parity}")
    out = {
        "metric": "Mrays/s of uneven batched ray queries",
        "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded world; seeded primary / shadow / AO rays)",
        "config": {"workload": "batched: 96 batches/step, sizes log-uniform 256..262144, 40% "
                               "coherent primary / 30% shadow / 30% cosine AO, tmax in "
                               "{32,128,512,1024}, 256x128x256 world, one HIP graph per step",
                   "world": list(world.dims), "rays": bs.total, "batches": bs.n,
                   "schedule": dict(sched, kernel="persistent ray-compacting (batches)")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(BATCHED_TRAFFIC_PROFILE,
                                                 "traffic_bytes_per_step"),
                     "traffic_source": BATCHED_TRAFFIC_PROFILE,
                     "kernel": "k_batches_persistent",
                     "kernel_ms": round(float(np.mean(kern_ms)), 4),
                     "algorithmic_bytes": int(bytes_per_step)},
        "cpu_baseline": base,
        "parity": parity if base is not None else None,
    }
    print(json.dumps(out), flush=True)


# The learned update's configuration: 16 gather rays per probe ending at the first hit in
# the radiance cache (16 hash-grid levels of 2^16 entries), trained online with one step of
# 4096 plain-traced bounce samples every fourth update. Measured on this workload: the error
# is flat across 1..8 updates per step, 2k..64k samples per step and 2^15..2^19 entries per
# level (1.80-1.86 %): the 16-ray gather quadrature sets it; the smaller table, batch and
# cadence make the training step's share of the chip smaller.
GI_LEARNED_RAYS = 16
GI_TRAIN_BATCH = 4096
GI_TRAIN_EVERY = 4
GI_CACHE_LOG2_TABLE = 16


def gi_probe_bench(ctx, dw, world, sun, steps, train_updates, threads):
    """GI probes over the world: a 32-voxel probe grid (64x8x64 on 2048x256x2048),
    sun + sky, 2-bounce diffuse. Plain N: N rays/probe, every sample traces gather, sun
    shadow, cosine bounce and its sun shadow (N = 256 is the quality reference; 32 and 16 are
    the cheaper plain settings). Learned: GI_LEARNED_RAYS rays/probe ending at the first hit
    in the radiance cache, trained online (GI_TRAIN_BATCH samples per step, one step every
    GI_TRAIN_EVERY updates, the steps inside the timed updates). Errors: mean relative SH-L1
    irradiance error of a fresh update against the FP64 4096-ray 2-bounce estimate of every
    live probe."""
    import numpy as np

    import oracle
    from voxelg_amd import gi, nrc
    from voxelg_amd._lib import VG_GI_SKIP_BURIED

    dims = world.dims
    grid = (dims[0] // 32, dims[1] // 32, dims[2] // 32)
    base = dict(grid=grid, spacing=32, epochs=1, update_div=1, bounces=2, flags=VG_GI_SKIP_BURIED)
    p256 = gi.gi_params(sun, rays_per_probe=256, **base)
    pl = gi.gi_params(sun, rays_per_probe=GI_LEARNED_RAYS, **base)
    stream = ctx.stream()

    kernel_ms = {}

    def span(fn, p, n, cache):
        a, b = ctx.event(), ctx.event()
        a.record(stream)
        for i in range(n):
            p["round"] = 1000 + i
            fn(p, i)
        if cache is not None:
            cache.join(stream)
        b.record(stream)
        stream.synchronize()
        return a.elapsed_ms(b) / n

    def timed(fn, p, n, cache=None, grid_obj=None, name=None):
        # n back-to-back updates (fn(p, i)) between two events on the stream; a learned
        # update's training step runs on the cache's own stream, so the end event waits for
        # the last one (vg_nrc_join): the span covers all the work of the n updates. With
        # grid_obj, the same n updates run once more with the grid recording each launch's
        # time (events between the launches, which lengthen the span: not the timed run).
        ms = span(fn, p, n, cache)
        if grid_obj is not None:
            grid_obj.set_timing(True)
            span(fn, p, n, cache)
            kernel_ms[name] = grid_obj.timing()
            grid_obj.set_timing(False)
        return ms

    from voxelg_amd import _lib as L

    def sched():
        return {"waves_per_cu": ctx.get_option(L.VG_OPT_PROBE_WAVES_PER_CU),
                "refill_below": ctx.get_option(L.VG_OPT_PROBE_REFILL),
                "ids_per_atomic": ctx.get_option(L.VG_OPT_PROBE_CHUNK),
                "transition_batch": ctx.get_option(L.VG_OPT_PROBE_TBATCH),
                "material_prefetch": ctx.get_option(L.VG_OPT_PROBE_MATSPEC)}

    def set_sched(sc):
        ctx.set_option(L.VG_OPT_PROBE_WAVES_PER_CU, sc["waves_per_cu"])
        ctx.set_option(L.VG_OPT_PROBE_REFILL, sc["refill_below"])
        ctx.set_option(L.VG_OPT_PROBE_CHUNK, sc["ids_per_atomic"])
        ctx.set_option(L.VG_OPT_PROBE_TBATCH, sc["transition_batch"])
        ctx.set_option(L.VG_OPT_PROBE_MATSPEC, sc["material_prefetch"])

    def train_on(i):
        return i % GI_TRAIN_EVERY == 0

    # plain settings: 256 (the quality reference) and the cheaper 32 and 16
    plain = {}
    for n in (256, 32, 16):
        p = p256 if n == 256 else gi.gi_params(sun, rays_per_probe=n, **base)
        gobj = gi.ProbeGrid(dw, p)
        rays_n, _ = gobj.update_counts(p)
        gobj.autotune(p)
        s = sched()
        ms = timed(lambda q, i: gobj.update(q, stream), p, steps, grid_obj=gobj,
                   name=f"plain{n}")
        plain[n] = {"grid": gobj, "params": p, "ms": ms, "sched": s, "rays": rays_n}
        log(f"gi plain {n} rays: {ms:.3f} ms/update ({s}), rays {rays_n}")
    plain_ms = plain[256]["ms"]

    cfg = nrc.default_config(log2_table=GI_CACHE_LOG2_TABLE)
    cfg["batch"] = GI_TRAIN_BATCH
    cache = nrc.RadianceCache(ctx, cfg)
    gl = gi.ProbeGrid(dw, pl)
    t0 = time.perf_counter()
    for r in range(train_updates):
        pl["round"] = r
        gl.update_learned(cache, pl, train=train_on(r), stream=stream)
    stream.synchronize()
    log(f"gi learned: {train_updates} training updates in {time.perf_counter() - t0:.2f} s, "
        f"loss {cache.stats()[1]:.5f}")
    # two candidate schedules: tuned on updates that each take a training step (the tuner's
    # pairs), and tuned on inference-only updates; the learned update (a step every
    # GI_TRAIN_EVERY updates) is timed under both and keeps the faster
    gl.autotune(pl, cache, train=True)  # (its repetitions train the cache further)
    sched_train = sched()
    gl.autotune(pl, cache, train=False)  # the inference-only update has its own schedule
    sched_frozen = sched()
    frozen_ms = timed(lambda p, i: gl.update_learned(cache, p, train=False, stream=stream), pl,
                      steps, cache, gl, "learned_inference_only")
    # an even number of updates, so the span holds whole training cadences
    lsteps = max(GI_TRAIN_EVERY, steps - steps % GI_TRAIN_EVERY)

    def learned_run(p, i):
        gl.update_learned(cache, p, train=train_on(i), stream=stream)

    cand = {}
    for name, sc in (("tuned_with_training", sched_train), ("tuned_inference_only", sched_frozen)):
        set_sched(sc)
        cand[name] = span(learned_run, pl, lsteps, cache)
    best = min(cand, key=cand.get)
    sched_learned = dict(sched_train if best == "tuned_with_training" else sched_frozen,
                         picked=best)
    set_sched(sched_learned)
    learned_ms = timed(learned_run, pl, lsteps, cache, gl, "learned")
    log(f"gi learned {GI_LEARNED_RAYS} rays: {learned_ms:.3f} ms/update with training "
        f"({sched_learned}; candidates {cand}), {frozen_ms:.3f} without ({sched_frozen})")
    pl["round"] = 5000
    raysl, nq, itl = gl.update_learned_counts(cache, pl, train=True, classes=True)
    log(f"gi learned: rays {raysl}, cache queries {nq}, walk classes {itl}; "
        f"launch ms {kernel_ms}")
    # roofline of the learned update's dominant launch, the path kernel, taken on the
    # inference-only update (a training update's launch also traces the step's targets): the
    # pyramid words its walks read (iteration classes) + per gather sample its probe origin,
    # direction, inverse and face byte (49 B) + per cache query its hit record, radiance and
    # list entry (17 + 32 B)
    from voxelg_amd._lib import algorithmic_bytes
    pl["round"] = 5001
    raysf, nqf, itf = gl.update_learned_counts(cache, pl, train=False, classes=True)
    path_bytes_frozen = algorithmic_bytes(itf) + 49 * raysf["gather"] + 49 * nqf
    path_ms = kernel_ms["learned_inference_only"]["paths"]
    path_gbs = path_bytes_frozen / (path_ms / 1e3) / 1e9

    # the probes' consumer: per-pixel indirect light over the headline 1080p frame's hits
    from voxelg_amd import workloads
    fw = workloads.LARGE
    _, fcam = fw.camera()
    fhits, _ = dw.trace_frame(fcam, fw.width, fw.height, fw.primary_tmax, fw.primary_flags,
                              fw.sun(), fw.shadow_tmax)
    find = ctx.alloc(fw.width * fw.height * 16)
    indirect_ms = timed(lambda p, i: gl.frame_indirect(p, fcam, fw.width, fw.height, fhits,
                                                        out=find, stream=stream), pl, steps)
    log(f"gi: per-pixel probe indirect at {fw.width}x{fw.height}: {indirect_ms:.4f} ms")

    # errors of fresh updates against the FP64 reference
    g256 = plain[256]["grid"]
    total = g256.total
    slots = np.arange(total)

    def fresh(grid_obj, fn):
        grid_obj.reset()
        fn()
        ctx.synchronize()
        return grid_obj.download()

    recs = {}
    for n, e in plain.items():
        e["params"]["round"] = 7000
        recs[f"plain{n}"] = fresh(e["grid"], lambda: e["grid"].update(e["params"]))
    pl["round"] = 7000
    recs["learned"] = fresh(gl, lambda: gl.update_learned(cache, pl, train=True))
    p1 = pl.copy()
    p1["bounces"] = 1
    recs["ambient"] = fresh(gl, lambda: gl.update(p1))
    live = recs["plain256"]["info"][:, 3] == 1
    origins = gi.probe_origins(np.zeros(total, gi.PROBE_DTYPE), p256, slots)[live]
    pref = p256.copy()
    pref["rays_per_probe"] = 4096
    pref["seed"] = 0xFEED
    t0 = time.perf_counter()
    ow = oracle.OracleWorld.view(world)
    ref = ow.gi_reference(pref, gi.directions(4096), gi.bounce_table(), gi.default_palette(),
                          slots[live], origins, threads=threads)
    ref_s = time.perf_counter() - t0

    def err(rec):
        return float(gi.irradiance_error(gi.sh_coefficients(rec)[live], ref).mean())

    errs = {k: err(v) for k, v in recs.items()}
    under = {f"plain{n}": e["ms"] for n, e in plain.items() if errs[f"plain{n}"] <= GI_ERR_LIMIT}
    fastest = min(under, key=under.get) if under else None
    ok = errs["learned"] <= GI_ERR_LIMIT
    for e in plain.values():
        e["grid"].close()
    out = {
        "value": round(learned_ms, 4), "unit": "ms/update", "higher_is_better": False,
        "err_limit": GI_ERR_LIMIT, "learned_meets_err_limit": bool(ok),
        "speedup_vs_plain256": round(plain_ms / learned_ms, 3) if ok else None,
        "target_speedup": 4.0,
        "fastest_plain_under_limit": fastest,
        "speedup_vs_fastest_plain_under_limit":
            round(under[fastest] / learned_ms, 3) if (ok and fastest) else None,
        "probes": int(total), "live_probes": int(live.sum()), "grid": list(grid), "spacing": 32,
        "learned": {"rays_per_probe": GI_LEARNED_RAYS, "train_batch": GI_TRAIN_BATCH,
                    "train_every": GI_TRAIN_EVERY, "log2_table": GI_CACHE_LOG2_TABLE,
                    "timed_updates": lsteps},
        "plain_ms": round(plain_ms, 4), "learned_ms": round(learned_ms, 4),
        "plain32_ms": round(plain[32]["ms"], 4), "plain16_ms": round(plain[16]["ms"], 4),
        "speedup": round(plain_ms / learned_ms, 3),
        "learned_inference_only_ms": round(frozen_ms, 4),
        "speedup_inference_only": round(plain_ms / frozen_ms, 3),
        "schedule_plain": plain[256]["sched"], "schedule_plain32": plain[32]["sched"],
        "schedule_plain16": plain[16]["sched"], "schedule_learned": sched_learned,
        "schedule_inference_only": sched_frozen,
        "learned_schedule_candidates_ms": {k: round(v, 4) for k, v in cand.items()},
        "err_plain256": round(errs["plain256"], 5),
        "err_plain32_2bounce": round(errs["plain32"], 5),
        "err_plain16_2bounce": round(errs["plain16"], 5),
        "err_learned": round(errs["learned"], 5),
        "err_ambient_tail": round(errs["ambient"], 5),
        "rays_plain": plain[256]["rays"], "rays_learned": raysl, "cache_queries": nq,
        "launch_ms": {k: {kk: round(vv, 4) if isinstance(vv, float) else vv
                          for kk, vv in v.items()} for k, v in kernel_ms.items()},
        "roofline": {"bound": "hbm", "achieved": round(path_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(path_gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(GI_TRAFFIC_PROFILE, "traffic_bytes_per_update"),
                     "traffic_source": GI_TRAFFIC_PROFILE,
                     "kernel": "k_probe_paths<false, 4, *, *> (the learned update's path "
                               "launch, an update without a training step)",
                     "kernel_ms": round(path_ms, 4), "algorithmic_bytes": int(path_bytes_frozen)},
        "indirect_1080p_ms": round(indirect_ms, 4),
        "train_steps": cache.stats()[0], "train_batch": cache.batch,
        "reference": f"FP64 4096 rays/probe 2-bounce, oracle on {threads} host threads, "
                     f"{ref_s:.1f} s",
    }
    log(f"gi: {out}")
    return out


def bench_gi_main(args):
    """One step = one learned GI probe update (GI_LEARNED_RAYS rays/probe, the cache trained
    online) on the 2048x256x2048 world; the plain 256-ray update is timed beside it."""
    from voxelg_amd import workloads
    from voxelg_amd.device import Context, device_count

    assert device_count() > 0, "bench.py needs a HIP device"
    wl = workloads.LARGE
    world = wl.world()
    sun = wl.sun()
    ctx = Context(0)
    dw = ctx.world(wl.dims).upload(world)
    g = gi_probe_bench(ctx, dw, world, sun, args.steps, args.gi_train_updates, gi_threads())
    out = {
        "metric": "GI probe ms at <=2% irradiance err vs plain",
        "value": g["learned_ms"], "unit": "ms/update", "n_gpus": 1, "steps": args.steps,
        "warmup": args.gi_train_updates, "ms_per_step": g["learned_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16",
        "data": "synthetic (seeded fBm terrain with caves, generated on the host)",
        "config": {"workload": "GI probes: 64x8x64 grid at 32-voxel spacing over the "
                               "2048x256x2048 world, sun + sky, 2-bounce diffuse; learned "
                               f"{GI_LEARNED_RAYS} rays/probe + hash-grid MLP cache trained "
                               f"online ({GI_TRAIN_BATCH} samples per step, one step every "
                               f"{GI_TRAIN_EVERY} updates) vs plain 256 rays/probe",
                   "world": list(wl.dims)},
        "gi": g,
        "cache_mfma_util": {"train": profile_value(COUNTERS_PROFILE, "mfma_util_train"),
                            "infer": profile_value(COUNTERS_PROFILE, "mfma_util_infer"),
                            "source": COUNTERS_PROFILE},
    }
    print(json.dumps(out), flush=True)


def bench_dynamic(args):
    """Dynamic world: the 256x128x256 world plus 4096 seeded sphere edits per frame (radius 2..12,
    half carve / half fill). One step = one frame: apply the edits on the device (incremental
    pyramid / uniform rebuild), mark the dirtied 8x8 blocks (primary and sun-shadow rays),
    retrace them at 1920x1080 (primary tmax 512 + shadow 480), a learned probe update
    (8x4x8 probes at spacing 32, 32 rays per probe, the cache training online), then the
    probes' per-pixel indirect light over the frame's hits (sample_probes). Every
    --check-every frames the incrementally maintained frame is compared bit for bit with the
    oracle's full frame of the same edited world (outside the timed region)."""
    import math

    import numpy as np

    import oracle
    import voxelg_amd as vg
    from voxelg_amd import gi, nrc, workloads
    from voxelg_amd._lib import VG_GI_SKIP_BURIED
    from voxelg_amd.device import Context, decode_hits, device_count, mask_words

    assert device_count() > 0, "bench.py needs a HIP device"
    wl = workloads.DYNAMIC
    world = wl.world()
    ref = oracle.OracleWorld.copy_of(world)
    ctx = Context(0)
    dw = ctx.world(wl.dims).upload(world)
    W, H = wl.width, wl.height
    _, cam = workloads.FrameWorkload("dynamic", wl.dims, W, H, 512.0, 0, True).camera()
    sun = vg.sun_direction(math.radians(35.0), 2.1)
    stream = ctx.stream()
    hits = ctx.alloc(W * H * 20)
    occ = ctx.alloc(W * H)
    dw.trace_frame(cam, W, H, 512.0, 0, sun, 480.0, hits, occ, stream)
    pgrid = (wl.dims[0] // 32, wl.dims[1] // 32, wl.dims[2] // 32)
    p = gi.gi_params(sun, grid=pgrid, spacing=32, rays_per_probe=32, epochs=1, update_div=1,
                     bounces=2, flags=VG_GI_SKIP_BURIED)
    probes = gi.ProbeGrid(dw, p)
    # the GI bench's cache size and batch (the world changes every frame: a step every frame)
    cfg = nrc.default_config(log2_table=GI_CACHE_LOG2_TABLE)
    cfg["batch"] = GI_TRAIN_BATCH
    cache = nrc.RadianceCache(ctx, cfg)
    # schedules tuned on the unedited world before the timed frames: the probe update's with
    # its training step (the tuning's updates train the cache too), the frame kernel's on the
    # full frame (the dirty-block retrace keeps 8x8 blocks and takes the rest)
    probes.autotune(p, cache, train=True)
    dw.autotune_frame(cam, W, H, 512.0, 0, sun, 480.0)
    n_frames = args.warmup + args.steps
    edits = [wl.edits(f, world) for f in range(n_frames)]
    phases = ("edits", "dirty", "retrace", "probes", "indirect")
    ind = ctx.alloc(W * H * 16)
    mask_buf = ctx.alloc(4 * mask_words(W, H))  # one block mask, reused every frame
    times = {k: [] for k in phases}
    dirty_frac = []
    checks = []
    wall = 0.0
    nbx, nby = (W + 7) // 8, (H + 7) // 8
    for f in range(n_frames):
        ev = [ctx.event() for _ in range(6)]
        t0 = time.perf_counter()
        ev[0].record(stream)
        dw.apply_edits(edits[f], stream)
        ev[1].record(stream)
        mask = dw.dirty_blocks(cam, W, H, sun, 480.0, mask=mask_buf, stream=stream)
        ev[2].record(stream)
        dw.trace_frame_blocks(cam, W, H, 512.0, 0, sun, 480.0, mask, hits, occ, stream)
        ev[3].record(stream)
        p["round"] = f
        probes.update_learned(cache, p, train=True, stream=stream)
        ev[4].record(stream)
        probes.frame_indirect(p, cam, W, H, hits, out=ind, stream=stream)
        ev[5].record(stream)
        if f == n_frames - 1:
            # the cache's training step overlaps the next frame; the last one is waited for
            cache.join(stream)
        stream.synchronize()
        dt = time.perf_counter() - t0
        if f >= args.warmup:
            wall += dt
            for k, name in enumerate(phases):
                times[name].append(ev[k].elapsed_ms(ev[k + 1]))
            m = mask.numpy(np.uint32)
            bits = np.unpackbits(m[:-1].view(np.uint8), bitorder="little")[:nbx * nby]
            dirty_frac.append(1.0 if m[-1] else float(bits.mean()))
        ref.apply_spheres(edits[f])
        if args.check_every > 0 and (f + 1) % args.check_every == 0:
            exp, exp_occ, _ = ref.frame(cam, W, H, 512.0, 0, sun, 480.0,
                                        threads=min(os.cpu_count() or 1, 16))
            ok = decode_hits(hits).tobytes() == exp.tobytes() and \
                np.array_equal(occ.numpy(np.uint8), exp_occ)
            checks.append({"frame": f, "bit_exact": bool(ok)})
            log(f"frame {f}: bit-exact check {'ok' if ok else 'FAILED'}")
            assert ok, f"frame {f}: incrementally maintained frame differs from the oracle"
    ms = {k: round(float(np.mean(v)), 4) for k, v in times.items()}
    ms_frame = wall / args.steps * 1e3
    log(f"dynamic: {ms_frame:.3f} ms/frame wall; phases {ms}; dirty fraction "
        f"{np.mean(dirty_frac):.3f}; cache steps {cache.stats()[0]}")
    out = {
        "metric": "dynamic-world frame ms (4096 sphere edits + incremental rebuild + dirty "
                  "retrace + learned probe update + per-pixel probe GI)",
        "value": round(ms_frame, 4), "unit": "ms/frame", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_frame, 4), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16",
        "data": "synthetic (seeded world, seeded sphere edits)",
        "config": {"workload": "dynamic: 256x128x256 world, 4096 sphere edits per frame (r 2..12, "
                               "50% carve / 50% fill), 1920x1080 primary + sun shadow retrace of "
                               "the dirty 8x8 blocks, learned probe update (8x4x8 probes, 32 "
                               f"rays, online training: {GI_TRAIN_BATCH} samples per frame, "
                               f"2^{GI_CACHE_LOG2_TABLE} entries per level), per-pixel probe "
                               "indirect light",
                   "world": list(wl.dims), "resolution": [W, H], "edits_per_frame": wl.n_edits},
        "phases_ms": ms, "dirty_block_fraction": round(float(np.mean(dirty_frac)), 4),
        "checks": checks,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
