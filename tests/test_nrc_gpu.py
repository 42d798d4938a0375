"""GPU parity of the radiance cache (bf16 MFMA inference / training, Adam) and of the learned
probe update against the FP64 oracle, and the online-training gain.

The oracle runs with bf16 emulation: it rounds parameters, layer inputs and layer gradients
to bf16 where the device holds them in bf16 and sums in double. Device and oracle then differ
by fp32 accumulation order (and the rare bf16 rounding flip it causes), so the tolerances
below are norm-wise: 2e-3 for predictions, 1e-2 for gradients.
"""
import math
import os

import numpy as np
import pytest

import oracle
from voxelg_amd import gi, nrc

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
SUN = (math.cos(math.radians(35.0)) * math.cos(2.1), math.sin(math.radians(35.0)),
       math.cos(math.radians(35.0)) * math.sin(2.1))


def case(msg):
    print(f"case: {msg}", flush=True)


def rand_queries(rng, n, hi=300.0):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return nrc.queries(rng.uniform(0.0, hi, (n, 3)), rng.integers(0, 7, n), d,
                       aux=np.arange(n))


def rand_params(rng, cache, scale_mlp=0.2, scale_grid=0.5):
    mlp = (rng.normal(size=cache.n_mlp) * scale_mlp).astype(np.float32)
    mlp[12288 + 3 * 64:] = 0.0
    grid = (rng.normal(size=cache.n_grid) * scale_grid).astype(np.float32)
    return mlp, grid


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def cache12(ctx):
    """16 levels of 2^12 entries: a grid small enough to compare whole gradients."""
    return nrc.RadianceCache(ctx, nrc.default_config(log2_table=12, batch=4096))


def test_infer_matches_oracle(cache12):
    rng = np.random.default_rng(11)
    mlp, grid = rand_params(rng, cache12)
    cache12.set_params(mlp, grid)
    for n in (1, 63, 3001):
        case(f"nrc infer n={n}")
        q = rand_queries(rng, n)
        y = cache12.infer(q)
        ref = oracle.NrcOracle(cache12.config, mlp, grid, emulate=True).forward(q)
        assert rel(y, ref) < 2e-3, rel(y, ref)
        assert np.abs(y - ref).max() < 2e-2 * np.abs(ref).max() + 1e-6
    exact = oracle.NrcOracle(cache12.config, mlp, grid).forward(q)
    assert rel(y, exact) < 5e-2  # bf16 versus fp64 throughout


def test_gradients_match_oracle(cache12):
    rng = np.random.default_rng(12)
    mlp, grid = rand_params(rng, cache12)
    cache12.set_params(mlp, grid)
    n = 1500  # not a multiple of the 256-sample tile
    q = rand_queries(rng, n, hi=120.0)
    t = rng.uniform(-0.5, 1.5, (n, 4)).astype(np.float32)
    t[:, 3] = 1.0
    t[::7, 3] = 0.0  # zero-weight samples contribute nothing
    case(f"nrc gradients n={n}")
    loss, gm, gg = cache12.gradients(q, t)
    lo, om, og = oracle.NrcOracle(cache12.config, mlp, grid, emulate=True).loss_grad(q, t)
    assert loss == pytest.approx(lo, rel=2e-3)
    for lyr in range(4):
        sl = slice(lyr * 4096, lyr * 4096 + (4096 if lyr < 3 else 1024))
        assert rel(gm[sl], om[sl]) < 1e-2, (lyr, rel(gm[sl], om[sl]))
    assert np.all(gm[12288 + 3 * 64:] == 0)
    assert rel(gg, og) < 1e-2, rel(gg, og)
    # entries no query touches stay exactly zero
    assert np.all(gg[og == 0] == 0)
    # a second call gives the same gradients bit for bit: the MLP sums are reduced in a fixed
    # order and the grid contributions are sorted by entry and reduced by key (no atomics)
    _, gm2, gg2 = cache12.gradients(q, t)
    np.testing.assert_array_equal(gm2, gm)
    np.testing.assert_array_equal(gg2, gg)


def test_train_step_is_adam_on_the_gradient(cache12):
    rng = np.random.default_rng(13)
    mlp, grid = rand_params(rng, cache12)
    cache12.set_params(mlp, grid)
    q = rand_queries(rng, 2048, hi=120.0)
    t = rng.uniform(0.0, 1.0, (2048, 4)).astype(np.float32)
    t[:, 3] = 1.0
    case("nrc train step")
    _, gm, gg = cache12.gradients(q, t)
    cache12.train_step(q, t)
    steps, loss = cache12.stats()
    assert steps == 1
    m1, g1 = cache12.params()
    c = cache12.config[0]
    for name, p0, g, got, lr in (("mlp", mlp, gm, m1, c["lr_mlp"]),
                                 ("grid", grid, gg, g1, c["lr_grid"])):
        p = p0.astype(np.float64)
        oracle.nrc_adam(p, np.zeros_like(p), np.zeros_like(p), g.astype(np.float64), float(lr),
                        float(c["beta1"]), float(c["beta2"]), float(c["eps"]), 1, True)
        live = g != 0
        bad = np.abs(got - p) > 1e-5 * np.abs(p) + 1e-7
        if name == "mlp":  # deterministic gradient: every weight follows Adam
            assert not bad.any()
        else:
            # the train step recomputes the grid gradient, whose atomics may sum in another
            # order: a gradient that nearly cancels can change sign or reach exactly zero in
            # one of the two evaluations (a move of up to 2 lr); only such entries differ
            assert bad.sum() <= 1e-3 * live.sum(), bad.sum()
            assert np.all(np.abs(g[bad]) < 1e-2 * np.abs(g).max())
            assert np.all(np.abs(got[bad] - p0[bad]) <= 2.0 * float(lr) * (1 + 1e-5))
    # the second step uses the gradient at the new parameters
    cache12.train_step(q, t)
    assert cache12.stats()[0] == 2


def test_non_finite_batch_changes_nothing(cache12):
    # one NaN (or infinite) target makes the backward pass non-finite: the step then changes
    # neither the grid nor the MLP (its gradient maximum, the grid's fixed-point scale, is
    # +inf), and the next finite batch trains as usual
    rng = np.random.default_rng(14)
    mlp, grid = rand_params(rng, cache12)
    q = rand_queries(rng, 1024, hi=120.0)
    t = rng.uniform(0.0, 1.0, (1024, 4)).astype(np.float32)
    t[:, 3] = 1.0
    for bad in (np.nan, np.inf):
        cache12.set_params(mlp, grid)
        tb = t.copy()
        tb[517, 1] = bad
        case(f"nrc train step with a {bad} target")
        cache12.train_step(q, tb)
        m1, g1 = cache12.params()
        np.testing.assert_array_equal(m1, mlp)
        np.testing.assert_array_equal(g1, grid)
        assert not np.isfinite(cache12.stats()[1])
    cache12.train_step(q, t)
    m2, g2 = cache12.params()
    assert np.isfinite(m2).all() and np.isfinite(g2).all() and not np.array_equal(g2, grid)


def probe_setup(small, grid=(8, 4, 8), rays=32):
    p = gi.gi_params(SUN, grid=grid, spacing=32, rays_per_probe=rays, epochs=1, update_div=1,
                     bounces=2, flags=1)
    return p, gi.directions(rays), gi.bounce_table(), gi.default_palette()


def test_learned_update_matches_oracle(ctx, small, small_world):
    cache = nrc.RadianceCache(ctx, nrc.default_config(log2_table=14, batch=2048))
    rng = np.random.default_rng(14)
    mlp, grid = rand_params(rng, cache, scale_mlp=0.15, scale_grid=0.3)
    cache.set_params(mlp, grid)
    p, dirs, bt, pal = probe_setup(small)
    pg = gi.ProbeGrid(small, p)
    ow = oracle.OracleWorld.view(small_world)
    ref = np.zeros((pg.total, 48))
    ev = oracle.NrcOracle(cache.config, mlp, grid, emulate=True)
    for r in (0, 1):
        p["round"] = r
        case(f"learned probe update round {r}")
        rays, nq = pg.update_learned_counts(cache, p, train=False)
        orays = ow.gi_update_learned(p, dirs, bt, pal, ev, ref, threads=THREADS)
        assert rays["gather"] == orays[0] and rays["shadow"] == orays[1] and rays["bounce"] == 0
        assert 100 < nq <= rays["gather"]
        got = pg.download().view(np.float32).reshape(pg.total, 48).astype(np.float64)
        for col in (20, 21, 22, 23, 44, 45, 46, 47):
            np.testing.assert_array_equal(got[:, col], ref[:, col])
        live = ref[:, 23] == 1
        assert live.sum() > 50
        for base in (0, 24):
            sh, rsh = got[live, base:base + 20], ref[live, base:base + 20]
            if np.abs(rsh).max() > 0:
                assert rel(sh, rsh) < 2e-3, (base, rel(sh, rsh))
    assert cache.stats()[0] == 0  # no training without the flag


def test_learned_update_with_material_prefetch(ctx, small, small_world):
    """The path launch's material prefetch (a schedule option) changes only when each walk
    step's material is loaded: the learned update gives the same records bit for bit with it
    on and off, round after round (counting runs never prefetch, so this uses the timed
    entry point)."""
    from voxelg_amd import _lib as L
    cfg = nrc.default_config(log2_table=14, batch=2048)
    a, b = nrc.RadianceCache(ctx, cfg), nrc.RadianceCache(ctx, cfg)
    mlp, grid = rand_params(np.random.default_rng(22), a, scale_mlp=0.15, scale_grid=0.3)
    a.set_params(mlp, grid)
    b.set_params(mlp, grid)
    p, _, _, _ = probe_setup(small)
    ga, gb = gi.ProbeGrid(small, p), gi.ProbeGrid(small, p)
    keep = ctx.get_option(L.VG_OPT_PROBE_MATSPEC)
    try:
        for r in range(3):
            p["round"] = r
            case(f"learned round {r}: material prefetch off / on")
            ctx.set_option(L.VG_OPT_PROBE_MATSPEC, 0)
            ga.update_learned(a, p, train=False)
            ctx.set_option(L.VG_OPT_PROBE_MATSPEC, 1)
            gb.update_learned(b, p, train=False)
            ctx.synchronize()
            assert ga.download().tobytes() == gb.download().tobytes(), f"round {r}"
    finally:
        ctx.set_option(L.VG_OPT_PROBE_MATSPEC, keep)


def test_training_targets_match_oracle(ctx, small, small_world):
    """A learned update trains on the previous learned update's hits: the first has none
    (every target weight 0), the second traces a bounce for each of its batch."""
    cache = nrc.RadianceCache(ctx, nrc.default_config(log2_table=14, batch=3000))
    p, dirs, bt, pal = probe_setup(small)
    pg = gi.ProbeGrid(small, p)
    p["round"] = 4
    case("learned update with training, no previous hits")
    rays0, nq0 = pg.update_learned_counts(cache, p, train=True)
    assert cache.stats()[0] == 1 and rays0["bounce"] == 0
    q, t = cache.last_batch()
    assert np.all(t == 0) and np.all(q["face"] == 6)
    p["round"] = 5
    case("learned update with training: targets")
    rays, nq = pg.update_learned_counts(cache, p, train=True)
    assert cache.stats()[0] == 2
    q, t = cache.last_batch()
    assert len(q) == 3000 and np.all(t[:, 3] == 1.0)
    assert rays["bounce"] == 3000  # every training sample traces its bounce ray
    ow = oracle.OracleWorld.view(small_world)
    ref = ow.gi_bounce_tails(p, bt, pal, q, q["aux"], threads=THREADS)
    np.testing.assert_allclose(t[:, :3], ref, rtol=2e-6, atol=1e-7)
    assert np.all(q["face"] < 6) and np.all(q["aux"] < 4096)
    # the batch is stratified over the previous update's nq0 hits: every hit is drawn
    keys = {(tuple(r["pos"]), int(r["face"])) for r in q}
    assert len(keys) == nq0


def test_online_training_gains_over_no_cache(ctx, small, small_world):
    """32 gather rays per probe: the trained cache's bounce tail versus the reference's
    ambient tail (the method without the network) and versus an untrained cache, all
    measured against the FP64 4096-ray 2-bounce estimate."""
    p, dirs, bt, pal = probe_setup(small)
    ow = oracle.OracleWorld.view(small_world)
    pg = gi.ProbeGrid(small, p)
    total = pg.total
    slots = np.arange(total)
    origins = gi.probe_origins(np.zeros(total, gi.PROBE_DTYPE), p, slots)
    pref = p.copy()
    pref["rays_per_probe"] = 4096
    ref = ow.gi_reference(pref, gi.directions(4096), bt, pal, slots, origins, threads=THREADS)

    def err():
        rec = pg.download()
        live = rec["info"][:, 3] == 1
        return float(gi.irradiance_error(gi.sh_coefficients(rec)[live], ref[live]).mean())

    cache = nrc.RadianceCache(ctx, nrc.default_config(batch=16384))
    case("untrained cache")
    pg.reset().update_learned(cache, p, train=False)
    e_untrained = err()
    p1 = p.copy()
    p1["bounces"] = 1
    pg.reset().update(p1)
    e_ambient = err()
    for r in range(80):
        p["round"] = r
        pg.update_learned(cache, p, train=True)
    ctx.synchronize()
    steps, loss = cache.stats()
    assert steps == 80
    p["round"] = 1000
    case("trained cache")
    pg.reset().update_learned(cache, p, train=False)
    e_trained = err()
    print(f"32-ray irradiance error: trained cache {e_trained:.4f}, ambient tail {e_ambient:.4f}, "
          f"untrained cache {e_untrained:.4f} (last loss {loss:.5f})", flush=True)
    assert e_trained < 0.75 * e_ambient
    assert e_trained < 0.75 * e_untrained


def test_learned_update_error_small_world(ctx, small, small_world):
    """The learned update on the 256x128x256 world (8x4x8 probes at spacing 32, 32 gather rays per
    probe, the cache trained online on 65 536 samples per update), against the FP64 4096-ray
    2-bounce estimate. The 2 % limit is a property of the ray count here: the plain update
    tracing both bounces with the same 32 directions is 2.9 % off (64 rays: 1.9 %), so the
    32-ray quadrature alone exceeds it on this world (on the 2048x256x2048 world bench.py
    gates both updates at 2 %). What the cache must do is lose nothing against tracing the
    bounce it replaces, and beat the ambient tail by far; plain 256 rays meets 2 %."""
    p, dirs, bt, pal = probe_setup(small)
    ow = oracle.OracleWorld.view(small_world)
    pg = gi.ProbeGrid(small, p)
    total = pg.total
    slots = np.arange(total)
    origins = gi.probe_origins(np.zeros(total, gi.PROBE_DTYPE), p, slots)
    pref = p.copy()
    pref["rays_per_probe"] = 4096
    ref = ow.gi_reference(pref, gi.directions(4096), bt, pal, slots, origins, threads=THREADS)

    def err(grid):
        rec = grid.download()
        live = rec["info"][:, 3] == 1
        assert live.sum() > 100
        return float(gi.irradiance_error(gi.sh_coefficients(rec)[live], ref[live]).mean())

    cache = nrc.RadianceCache(ctx)  # default configuration: 65 536 samples per update
    assert cache.batch == 65536
    case("learned update: 150 online-training updates")
    for r in range(150):
        p["round"] = r
        pg.update_learned(cache, p, train=True)
    p["round"] = 7000
    pg.reset().update_learned(cache, p, train=True)
    ctx.synchronize()
    e_learned = err(pg)

    def plain(rays, bounces):
        q = gi.gi_params(SUN, grid=(8, 4, 8), spacing=32, rays_per_probe=rays, epochs=1,
                         update_div=1, bounces=bounces, flags=1)
        q["round"] = 7000
        g = gi.ProbeGrid(small, q)
        g.update(q)
        ctx.synchronize()
        return err(g)

    e32, e_ambient, e256 = plain(32, 2), plain(32, 1), plain(256, 2)
    print(f"256x128x256 world: learned 32-ray {e_learned:.4f}, plain 32-ray 2-bounce {e32:.4f}, "
          f"32-ray ambient tail {e_ambient:.4f}, plain 256-ray {e256:.4f}", flush=True)
    assert e_learned <= 1.02 * e32, (e_learned, e32)
    assert e_learned < 0.5 * e_ambient
    assert e256 <= 0.02


def test_probe_updates_replay_from_graphs(ctx, small, small_world):
    """Plain and learned probe updates captured into HIP graphs: every replay (eager work
    between them) gives the records of the same update run eagerly from the same state; a
    captured training update trains the cache once per replay, its records coming from the
    parameters before its own step."""
    from voxelg_amd.device import Graph
    p, dirs, bt, pal = probe_setup(small)
    p["round"] = 3
    rng = np.random.default_rng(21)
    stream = ctx.stream()
    other = oracle.make_rays(rng.uniform(0, 200, (20000, 3)), rng.normal(size=(20000, 3)))

    def same_records(a, b, what):
        assert a.tobytes() == b.tobytes(), what

    # plain 2-bounce update
    pe, pgr = gi.ProbeGrid(small, p), gi.ProbeGrid(small, p)
    gp = Graph(ctx, stream, lambda s: pgr.update(p, stream=s))
    for rep in range(3):
        case(f"plain update graph replay {rep}")
        pe.reset()
        pgr.reset()
        ctx.synchronize()
        pe.update(p)
        small.trace_rays(ctx.rays(other))  # eager work between replays
        ctx.synchronize()
        gp.launch(stream)
        stream.synchronize()
        same_records(pgr.download(), pe.download(), f"plain replay {rep}")
    # learned, inference only
    cache = nrc.RadianceCache(ctx, nrc.default_config(batch=4096))
    mlp, grid = rand_params(rng, cache, 0.05, 0.05)
    cache.set_params(mlp, grid)
    le, lg = gi.ProbeGrid(small, p), gi.ProbeGrid(small, p)
    gl = Graph(ctx, stream, lambda s: lg.update_learned(cache, p, train=False, stream=s))
    for rep in range(3):
        case(f"learned update graph replay {rep}")
        le.reset()
        lg.reset()
        ctx.synchronize()
        le.update_learned(cache, p, train=False)
        small.trace_rays(ctx.rays(other))
        ctx.synchronize()
        gl.launch(stream)
        stream.synchronize()
        same_records(lg.download(), le.download(), f"learned replay {rep}")
    # learned with training: the step is joined into the graph
    train_cache = nrc.RadianceCache(ctx, nrc.default_config(batch=4096))
    train_cache.set_params(mlp, grid)
    twin_cache = nrc.RadianceCache(ctx, nrc.default_config(batch=4096))
    lt, twin = gi.ProbeGrid(small, p), gi.ProbeGrid(small, p)
    lt.update_learned(train_cache, p, train=True)  # a first hit list to draw targets from
    train_cache.join()
    ctx.synchronize()
    steps0 = train_cache.stats()[0]
    gt = Graph(ctx, stream, lambda s: lt.update_learned(train_cache, p, train=True, stream=s))
    for rep in range(3):
        case(f"training update graph replay {rep}")
        m0, g0 = train_cache.params()
        twin_cache.set_params(m0, g0)
        twin.upload(lt.download())
        ctx.synchronize()
        gt.launch(stream)
        stream.synchronize()
        twin.update_learned(twin_cache, p, train=False)
        ctx.synchronize()
        same_records(lt.download(), twin.download(), f"training replay {rep}")
        m1, g1 = train_cache.params()
        assert train_cache.stats()[0] == steps0 + rep + 1
        assert np.isfinite(m1).all() and np.isfinite(g1).all() and not np.array_equal(g1, g0)
