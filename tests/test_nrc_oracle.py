"""Pins of the radiance-cache oracle (oracle/oracle_nrc.c) and of the learned probe update's
oracle (oracle/oracle_gi.c), independent of the device.

What fixes each function:
  bf16 rounding         its definition (round to nearest, ties to even) on exact tie cases
  hash                  the integer formula x*1 ^ y*2654435761 ^ z*805459861 mod 2^k
  scales                closed form b^l / coarse, endpoints 1/coarse and 1/fine
  encoding              partition of unity; a grid holding a linear function of the corner
                        coordinate interpolates to that function; the extra inputs' slots
  loss gradient         central finite differences on every layer and on grid entries
  Adam                  a hand-stepped recurrence; the sparse rule
  learned GI update     a zero cache == the 1-bounce update with no tail; a constant cache
                        == the 1-bounce update whose ambient tail is that constant
  bounce tail           == the bounce part of the plain 2-bounce gather sample
"""
import math

import numpy as np
import pytest

import oracle
from voxelg_amd import gi

SUN35 = (math.cos(math.radians(35.0)) * math.cos(2.1), math.sin(math.radians(35.0)),
         math.cos(math.radians(35.0)) * math.sin(2.1))


def small_config(levels=4, log2_table=8, coarse=16.0, fine=2.0):
    c = np.zeros(1, oracle.NRC_CONFIG_DTYPE)
    c["levels"] = levels
    c["log2_table"] = log2_table
    c["scale"][0, :levels] = oracle.nrc_scales(levels, coarse, fine)
    return c


def random_queries(rng, n, lo=0.0, hi=40.0):
    q = np.zeros(n, oracle.NRC_QUERY_DTYPE)
    q["pos"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    q["face"] = rng.integers(0, 7, n)
    d = rng.normal(size=(n, 3))
    q["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return q


def random_params(rng, cfg, scale_mlp=0.25, scale_grid=0.5):
    n_grid = int(cfg["levels"][0]) * (1 << int(cfg["log2_table"][0])) * 2
    mlp = (rng.normal(size=oracle.NRC_MLP_PARAMS) * scale_mlp).astype(np.float32)
    mlp[12288 + 3 * 64:] = 0.0  # W3 rows 3..15 stay zero
    grid = (rng.normal(size=n_grid) * scale_grid).astype(np.float32)
    return mlp, grid


def test_bf16_round_definition():
    # exactly representable values are kept
    for v in (0.0, 1.0, -2.5, 1.0 + 2 ** -7, 2.0 ** 127, 1.5 * 2.0 ** -126):
        assert oracle.bf16_round(v) == np.float32(v)
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 (mantissa ...0 even)
    assert oracle.bf16_round(1.0 + 2 ** -8) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    assert oracle.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    # just above a tie rounds up
    assert oracle.bf16_round(1.0 + 2 ** -8 + 2 ** -20) == 1.0 + 2 ** -7
    # bf16 has fp32's exponent range
    assert oracle.bf16_round(1e-30) == pytest.approx(1e-30, rel=2 ** -8)


def test_hash_closed_form():
    P1, P2 = 2654435761, 805459861
    rng = np.random.default_rng(3)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(-5000, 5000, 3))
        k = int(rng.integers(4, 23))
        want = ((x & 0xFFFFFFFF) ^ ((y * P1) & 0xFFFFFFFF) ^ ((z * P2) & 0xFFFFFFFF)) & ((1 << k) - 1)
        assert oracle.nrc_hash(x, y, z, k) == want


def test_scales_closed_form():
    s = oracle.nrc_scales(16, 128.0, 1.0)
    assert s[0] == np.float32(1.0 / 128.0) and s[-1] == np.float32(1.0)
    b = 128.0 ** (1.0 / 15.0)
    np.testing.assert_allclose(s, [b ** l / 128.0 for l in range(16)], rtol=1e-7)
    assert np.all(np.diff(s) > 0)


def test_encoding_partition_of_unity_and_extra_inputs():
    cfg = small_config()
    rng = np.random.default_rng(4)
    grid = np.ones(4 * 256 * 2, np.float32)
    grid[1::2] = -0.5  # feature 1 everywhere
    mlp = np.zeros(oracle.NRC_MLP_PARAMS, np.float32)
    o = oracle.NrcOracle(cfg, mlp, grid)
    for q in random_queries(rng, 20):
        x = o.encode(q)
        np.testing.assert_allclose(x[0:8:2], 1.0, rtol=0, atol=1e-12)
        np.testing.assert_allclose(x[1:8:2], -0.5, rtol=0, atol=1e-12)
        assert np.all(x[8:32] == 0)  # levels beyond cfg.levels
        face = int(q["face"])
        onehot = np.zeros(6)
        if face < 6:
            onehot[face] = 1
        assert np.array_equal(x[32:38], onehot)
        assert np.array_equal(x[38:41], q["dir"].astype(np.float64))
        assert x[41] == 1.0 and np.all(x[42:] == 0)


def test_encoding_interpolates_linear_functions():
    """Level 0 of a coarse grid, corners of one cell region hashed without collision: a grid
    holding (cx, cz) at corner (cx, cy, cz) interpolates to the scaled position."""
    cfg = small_config(levels=1, log2_table=16, coarse=8.0, fine=8.0)
    s = float(cfg["scale"][0, 0])
    grid = np.zeros((1 << 16) * 2, np.float32)
    seen = {}
    for cx in range(0, 6):
        for cy in range(0, 6):
            for cz in range(0, 6):
                h = oracle.nrc_hash(cx, cy, cz, 16)
                assert h not in seen, "collision in the test region"
                seen[h] = True
                grid[2 * h] = cx
                grid[2 * h + 1] = cz
    o = oracle.NrcOracle(cfg, np.zeros(oracle.NRC_MLP_PARAMS, np.float32), grid)
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.uniform(0, 4.99 / s, 3).astype(np.float32)
        q = np.zeros(1, oracle.NRC_QUERY_DTYPE)
        q["pos"] = p
        x = o.encode(q)
        assert x[0] == pytest.approx(float(np.float32(p[0] * np.float32(s))), abs=1e-6)
        assert x[1] == pytest.approx(float(np.float32(p[2] * np.float32(s))), abs=1e-6)


def fd_grad(o, q, t, which, i):
    """Central difference of the loss in parameter i of `which` (the actual fp32 step)."""
    arr = o.mlp if which == "mlp" else o.grid
    a0 = float(arr[i])
    h = max(abs(a0) * 1e-3, 1e-3)
    arr[i] = np.float32(a0 + h)
    up = float(arr[i])
    lp = o.loss(q, t)
    arr[i] = np.float32(a0 - h)
    dn = float(arr[i])
    lm = o.loss(q, t)
    arr[i] = np.float32(a0)
    return (lp - lm) / (up - dn)


def test_loss_gradient_matches_finite_differences():
    cfg = small_config()
    rng = np.random.default_rng(6)
    mlp, grid = random_params(rng, cfg)
    q = random_queries(rng, 7)
    t = rng.uniform(-0.5, 1.0, (7, 4)).astype(np.float32)
    t[:, 3] = rng.uniform(0.5, 2.0, 7)
    o = oracle.NrcOracle(cfg, mlp, grid)
    loss, gm, gg = o.loss_grad(q, t)
    assert loss == pytest.approx(o.loss(q, t), rel=1e-12)
    # every layer, several weights each (W3: the three live rows), incl. the bias column 41
    picks = []
    for lyr in range(3):
        picks += [lyr * 4096 + int(o_) * 64 + int(i_) for o_, i_ in
                  zip(rng.integers(0, 64, 8), rng.integers(0, 42, 8))]
        picks.append(lyr * 4096 + 5 * 64 + 41)
    picks += [12288 + r * 64 + int(i_) for r in range(3) for i_ in rng.integers(0, 64, 4)]
    for i in picks:
        fd = fd_grad(o, q, t, "mlp", i)
        assert gm[i] == pytest.approx(fd, rel=2e-4, abs=1e-8), f"mlp[{i}]"
    # W3 padding rows and dead input columns carry no gradient
    assert np.all(gm[12288 + 3 * 64:] == 0)
    assert np.all(gm[0:4096].reshape(64, 64)[:, 42:] == 0)
    nz = np.nonzero(gg)[0]
    assert len(nz) > 100
    for i in rng.choice(nz, 25, replace=False):
        fd = fd_grad(o, q, t, "grid", int(i))
        assert gg[i] == pytest.approx(fd, rel=2e-4, abs=1e-8), f"grid[{i}]"
    # an untouched grid entry has zero gradient
    z = np.setdiff1d(np.arange(len(gg)), nz)[0]
    assert fd_grad(o, q, t, "grid", int(z)) == 0.0


def test_emulated_bf16_is_close_to_exact():
    cfg = small_config()
    rng = np.random.default_rng(7)
    mlp, grid = random_params(rng, cfg)
    q = random_queries(rng, 64)
    y = oracle.NrcOracle(cfg, mlp, grid).forward(q)
    yb = oracle.NrcOracle(cfg, mlp, grid, emulate=True).forward(q)
    rel = np.linalg.norm(yb - y) / np.linalg.norm(y)
    assert 0 < rel < 3e-2


def test_adam_recurrence_by_hand():
    p = np.array([1.0, -2.0, 0.5])
    m = np.zeros(3)
    v = np.zeros(3)
    g1 = np.array([0.1, 0.0, -0.3])
    g2 = np.array([0.2, 0.0, 0.1])
    lr, b1, b2, eps = 0.01, 0.9, 0.99, 1e-10
    oracle.nrc_adam(p, m, v, g1, lr, b1, b2, eps, 1, True)
    oracle.nrc_adam(p, m, v, g2, lr, b1, b2, eps, 2, True)
    # by hand, entries 0 and 2 (entry 1 has zero gradients: sparse rule leaves it)
    for i, want_p0 in ((0, 1.0), (2, 0.5)):
        mm = vv = 0.0
        pp = want_p0
        for step, g in ((1, g1[i]), (2, g2[i])):
            mm = b1 * mm + (1 - b1) * g
            vv = b2 * vv + (1 - b2) * g * g
            pp -= lr * (mm / (1 - b1 ** step)) / (math.sqrt(vv / (1 - b2 ** step)) + eps)
        assert p[i] == pytest.approx(pp, rel=1e-14)
    assert p[1] == -2.0 and m[1] == 0.0 and v[1] == 0.0
    # the first step of Adam moves every live parameter by lr against its gradient's sign
    p2 = np.array([0.0, 0.0])
    oracle.nrc_adam(p2, np.zeros(2), np.zeros(2), np.array([3.0, -1e-6]), lr, b1, b2, 0.0, 1, True)
    np.testing.assert_allclose(p2, [-lr, lr], rtol=1e-12)


def _gi_setup(dims=(128, 128, 128), bounces=1, **kw):
    w = oracle.OracleWorld(dims).generate(0x5EED_0001)
    w.rebuild()
    p = gi.gi_params(SUN35, grid=(4, 4, 4), spacing=32, rays_per_probe=32, epochs=1,
                     update_div=1, bounces=bounces, flags=1, **kw)
    dirs = gi.directions(32)
    return w, p, dirs, gi.bounce_table(), gi.default_palette()


def constant_cache(k):
    """MLP whose output is the constant k (rgb) for every input: the bias input 41 drives
    one hidden unit through all three layers."""
    cfg = small_config()
    mlp = np.zeros(oracle.NRC_MLP_PARAMS, np.float32)
    mlp[0 * 4096 + 0 * 64 + 41] = 1.0
    mlp[1 * 4096 + 0] = 1.0
    mlp[2 * 4096 + 0] = 1.0
    for c in range(3):
        mlp[12288 + c * 64 + 0] = k[c]
    grid = np.zeros(4 * 256 * 2, np.float32)
    return oracle.NrcOracle(cfg, mlp, grid)


def test_learned_update_zero_cache_is_no_tail():
    w, p, dirs, bt, pal = _gi_setup(ambient_tail=0.0)
    a = np.zeros((64, 48))
    b = np.zeros((64, 48))
    w.gi_update(p, dirs, bt, pal, a)
    w.gi_update_learned(p, dirs, bt, pal, constant_cache([0.0, 0.0, 0.0]), b)
    assert (a[:, 23] == 1).sum() > 8
    np.testing.assert_array_equal(a, b)


def test_learned_update_constant_cache_is_constant_tail():
    k = np.array([0.25, 0.5, 0.125], np.float32)
    w, p, dirs, bt, pal = _gi_setup(ambient_tail=1.0)
    p["ambient"] = k
    a = np.zeros((64, 48))
    b = np.zeros((64, 48))
    w.gi_update(p, dirs, bt, pal, a)
    w.gi_update_learned(p, dirs, bt, pal, constant_cache(k), b)
    np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-15)
    # a negative cache output is clamped to zero
    a = np.zeros((64, 48))
    c = np.zeros((64, 48))
    p0 = p.copy()
    p0["ambient_tail"] = 0.0
    w.gi_update(p0, dirs, bt, pal, a)
    w.gi_update_learned(p0, dirs, bt, pal, constant_cache(-k), c)
    np.testing.assert_allclose(c, a, rtol=1e-13, atol=1e-15)


def test_bounce_tail_is_the_two_bounce_sample_tail():
    """radiance(2-bounce) = albedo * (direct + tail) with tail = orc_gi_bounce_tail at the
    gather hit, and radiance(no tail) = albedo * direct, sample by sample."""
    w, p2, dirs, bt, pal = _gi_setup(bounces=2, seed=11, round=3)
    p0 = p2.copy()
    p0["bounces"] = 1
    p0["ambient_tail"] = 0.0
    origins = gi.probe_origins(np.zeros(64, gi.PROBE_DTYPE), p2, np.arange(64))
    checked = 0
    for probe in range(0, 64, 3):
        o = origins[probe]
        if w.material_at(*np.floor(o).astype(int)) != 0:
            continue
        for k in range(32):
            d = dirs[k, :3]
            r2, f2, t2, _ = w.gi_sample(p2, bt, pal, o, d, probe, k)
            r0, f0, t0, _ = w.gi_sample(p0, bt, pal, o, d, probe, k)
            assert f2 == f0 and t2 == t0
            if f2 == oracle.FACE_NONE:
                np.testing.assert_array_equal(r2, r0)
                continue
            q = np.zeros(1, oracle.NRC_QUERY_DTYPE)
            q["pos"] = (o.astype(np.float32) + d.astype(np.float32) * np.float32(t2)).astype(np.float32)
            q["face"] = f2
            bidx = oracle.gi_bounce_index(int(p2["seed"][0]), int(p2["round"][0]), probe, k)
            tail = w.gi_bounce_tails(p2, bt, pal, q, [bidx], threads=1)[0]
            hit = w.trace_dir(o, d, float(p2["gi_dist"][0]), oracle.SKIP_TRANSPARENT)[1]
            alb = pal[int(hit["material"])][:3].astype(np.float64)
            np.testing.assert_allclose(r2[:3], r0[:3] + alb * tail, rtol=1e-12, atol=1e-15)
            checked += 1
    assert checked > 50
