"""The GI probe oracle pinned without the reference's shaders (which cannot run here):
closed forms of the update (constant sky, sealed boxes), sky_color where the reference
formula reduces to its constants, the direction tables against their definitions, the
relocation and fold recurrences stepped by hand, and agreement of the oracle's two
estimators on the same samples."""
import math

import numpy as np
import pytest

import oracle
from voxelg_amd import gi
from voxelg_amd._lib import GI_PARAMS_DTYPE, PROBE_DTYPE

SUN35 = (math.cos(math.radians(35.0)) * math.cos(0.7), math.sin(math.radians(35.0)),
         math.cos(math.radians(35.0)) * math.sin(0.7))


def test_params_layout_matches_oracle():
    assert GI_PARAMS_DTYPE == oracle.GI_PARAMS_DTYPE
    assert GI_PARAMS_DTYPE.itemsize == 152 and PROBE_DTYPE.itemsize == 192


def test_params_default_closed_forms():
    p = gi.gi_params(SUN35)[0]
    # the reference's constants (src/world_dims.rs:44-48, gi_probes.wgsl:248-272, rt_gi.wgsl:12-14)
    assert tuple(p["grid"]) == (64, 32, 64) and p["spacing"] == 8
    assert p["rays_per_probe"] == 8 and p["epochs"] == 8 and p["update_div"] == 8
    assert p["bounces"] == 1 and p["gi_dist"] == np.float32(40) and p["gi_sky"] == np.float32(0.32)
    assert p["hysteresis"] == np.float32(0.6) and p["shadow_tmax"] == np.float32(480)
    # a 35-degree sun is past every smoothstep: full day, neutral midday sun, no dusk wash
    assert p["day_t"] == 1.0 and p["dusk_w"] == 0.0
    np.testing.assert_array_equal(p["sun_color"], np.float32([1.10, 1.02, 0.92]))
    np.testing.assert_array_equal(p["ambient"], np.float32([0.30, 0.42, 0.58]))
    np.testing.assert_array_equal(p["sky_zenith"], np.float32([0.22, 0.50, 0.95]))
    np.testing.assert_array_equal(p["sky_horizon"], np.float32([0.78, 0.86, 0.95]))
    assert abs(p["halo"] - 0.45) < 1e-7
    assert abs(np.linalg.norm(p["sun"]) - 1.0) < 1e-6
    assert abs(np.linalg.norm(p["sun_xz"]) - 1.0) < 1e-6
    # a sun on the horizon: sun_intensity = smoothstep(-0.05, 0.10, 0) = t^2 (3 - 2t), t = 1/3
    q = gi.gi_params((1.0, 0.0, 0.0))[0]
    assert abs(q["day_t"] - 7.0 / 27.0) < 1e-6
    # and sun_color at h = 0 is the ember stop scaled by that intensity
    np.testing.assert_allclose(q["sun_color"], np.array([1.55, 0.52, 0.16]) * 7.0 / 27.0, rtol=1e-6)


def test_fibonacci_directions():
    n = 256
    d = gi.directions(n).astype(np.float64)
    assert np.all(d[:, 3] == 0)
    np.testing.assert_allclose(np.linalg.norm(d[:, :3], axis=1), 1.0, atol=1e-6)
    # z = 1 - 2 (i + 0.5) / n: equally spaced heights, symmetric about 0
    np.testing.assert_allclose(d[:, 1], 1.0 - 2.0 * (np.arange(n) + 0.5) / n, atol=1e-7)
    assert abs(d[:, 1].sum()) < 1e-5
    # consecutive points turn by the golden angle about the y axis
    phi = np.arctan2(d[:, 2], d[:, 0])
    step = np.mod(np.diff(phi), 2 * np.pi)
    np.testing.assert_allclose(step, np.mod(np.pi * (3 - math.sqrt(5)), 2 * np.pi), atol=1e-4)


def test_bounce_table_is_cosine_weighted():
    t = gi.bounce_table().astype(np.float64)
    assert t.shape == (4096, 4)
    np.testing.assert_allclose(np.linalg.norm(t[:, :3], axis=1), 1.0, atol=1e-6)
    assert np.all(t[:, 2] >= 0)
    # cosine-weighted hemisphere moments: E[cos] = 2/3, E[cos^2] = 1/2, E[t] = E[b] = 0
    assert abs(t[:, 2].mean() - 2.0 / 3.0) < 1e-3
    assert abs((t[:, 2] ** 2).mean() - 0.5) < 1e-3
    assert abs(t[:, 0].mean()) < 1e-3 and abs(t[:, 1].mean()) < 1e-3


def test_sky_color_special_directions():
    p = gi.gi_params(SUN35)
    s = p[0]["sun"].astype(np.float64)
    # straight down: horizon colour, no halo (sun below the ray)
    np.testing.assert_allclose(oracle.gi_sky(p, (0, -1, 0)), [0.78, 0.86, 0.95], atol=1e-7)
    # straight up: zenith + halo (1, .78, .52) * s.y^6 * 0.45
    h = s[1] ** 6 * 0.45
    np.testing.assert_allclose(oracle.gi_sky(p, (0, 1, 0)),
                               np.array([0.22, 0.50, 0.95]) + h * np.array([1, 0.78, 0.52]),
                               atol=1e-6)
    # into the sun: the disc, col * 0.1 + (2.4, 2.0, 1.5) * 0.9 with col = blend + halo 0.45
    zf = (lambda t: t * t * (3 - 2 * t))(min(1.0, s[1] / 0.55))
    base = np.array([0.78, 0.86, 0.95]) * (1 - zf) + np.array([0.22, 0.50, 0.95]) * zf
    col = base + 0.45 * np.array([1, 0.78, 0.52])
    np.testing.assert_allclose(oracle.gi_sky(p, s.astype(np.float32)),
                               col * 0.1 + 0.9 * np.array([2.4, 2.0, 1.5]), atol=1e-5)


def _tables(params):
    p = params[0]
    n = int(p["rays_per_probe"]) * int(p["epochs"])
    return gi.directions(n), gi.bounce_table(), gi.default_palette()


def _probes(params):
    g = params[0]["grid"]
    return np.zeros((int(g[0]) * int(g[1]) * int(g[2]), oracle.PROBE_FIELDS), np.float64)


def test_constant_sky_closed_form():
    """Empty world, sky made constant: every gather ray escapes with L = sky * GI_SKY, so
    c0 = 4 pi L 0.282095 exactly, the linear terms vanish on the symmetric set, and the
    sampler's irradiance/pi is L (gi_probes.wgsl:72-95)."""
    w = oracle.OracleWorld((64, 64, 64))
    L = np.array([0.5, 0.7, 0.9])
    p = gi.gi_params((0.0, -1.0, 0.0), grid=(2, 2, 2), spacing=32, rays_per_probe=128,
                     epochs=1, update_div=1, bounces=2)
    p["sky_horizon"] = L
    p["sky_zenith"] = L
    p["halo"] = 0.0
    dirs, bt, pal = _tables(p)
    assert np.max(dirs[:, :3] @ p[0]["sun"]) < 0.9985  # no direction in the sun disc
    pr = _probes(p)
    rays = w.gi_update(p, dirs, bt, pal, pr)
    assert rays == [8 * 128, 0, 0]
    Lg = L.astype(np.float32).astype(np.float64) * np.float64(np.float32(0.32))
    dsum = dirs[:, :3].astype(np.float64).sum(axis=0)  # the set's linear moments
    for c in range(3):
        np.testing.assert_allclose(pr[:, 4 * c], 4 * np.pi * Lg[c] * 0.282095, rtol=1e-12)
        lin = (4 * np.pi / 128) * Lg[c] * 0.488603 * dsum[[1, 2, 0]]
        np.testing.assert_allclose(pr[:, 4 * c + 1:4 * c + 4], np.broadcast_to(lin, (8, 3)),
                                   rtol=1e-9, atol=1e-15)
        assert np.all(np.abs(lin) < 5e-3 * pr[0, 4 * c])  # nearly symmetric set
    # visible distance is GI_DIST on every ray
    np.testing.assert_allclose(pr[:, 12], 4 * np.pi * 40.0 * 0.282095, rtol=1e-12)
    e = gi.sh_irradiance(pr[:, :12])  # = L up to the set's small linear moments
    np.testing.assert_allclose(e, np.broadcast_to(Lg, e.shape), rtol=2e-3)
    assert np.all(pr[:, 23] == 1.0)  # info.w: age 1


def _sealed_box(dims=64, lo=14, hi=34, mat=4):
    w = oracle.OracleWorld((dims, dims, dims))
    for x in range(lo, hi + 1):
        for y in range(lo, hi + 1):
            for z in range(lo, hi + 1):
                if x in (lo, hi) or y in (lo, hi) or z in (lo, hi):
                    w.set_voxel(x, y, z, mat)
    return w


@pytest.mark.parametrize("bounces", [1, 2])
def test_sealed_box_closed_form(bounces):
    """A probe inside a closed stone box: every gather ray hits a wall, no wall sees the
    sun, so L = a * ambient * tail (one bounce) or a * (a * ambient * tail) (the bounce ray
    hits another wall, also unlit: the box's inner diagonal, 19 sqrt(3), is within GI_DIST)."""
    w = _sealed_box()
    p = gi.gi_params(SUN35, grid=(1, 1, 1), spacing=48, rays_per_probe=256, epochs=1,
                     update_div=1, bounces=bounces)
    dirs, bt, pal = _tables(p)
    pr = _probes(p)
    rays = w.gi_update(p, dirs, bt, pal, pr)
    a = pal[4, :3].astype(np.float64)
    amb = p[0]["ambient"].astype(np.float64) * np.float64(p[0]["ambient_tail"])
    L = a * amb if bounces == 1 else a * (a * amb)
    for c in range(3):
        np.testing.assert_allclose(pr[0, 4 * c], 4 * np.pi * L[c] * 0.282095, rtol=1e-12)
    assert rays[0] == 256
    assert rays[2] == (256 if bounces == 2 else 0)
    # every sunward wall's shadow ray is blocked by the box
    assert rays[1] > 0
    # the probe at 24 sees the inner wall faces 9 and 10 voxels away along the axes
    mean_d = pr[0, 12] / (4 * np.pi * 0.282095)
    assert 9.0 < mean_d < 10.0 * math.sqrt(3) + 1


def test_relocation_recurrence_by_hand():
    """A probe half a voxel above a floor is pushed up by min(1.2 - near_t, 0.35) along the
    floor normal per update (gi_probes.wgsl:381-392), in fp32."""
    w = oracle.OracleWorld((64, 64, 64))
    for x in range(64):
        for z in range(64):
            w.set_voxel(x, 23, z, 4)  # top face at y = 24
    p = gi.gi_params(SUN35, grid=(1, 1, 1), spacing=49, rays_per_probe=64, epochs=1,
                     update_div=1, bounces=1)
    dirs, bt, pal = _tables(p)
    pr = _probes(p)
    f32 = np.float32
    off = f32(0.0)
    dmin = -dirs[:, 1].min()  # the steepest downward direction
    for step in range(4):
        y = f32(f32(f32(0) + f32(0.5)) * f32(49)) + off  # 24.5 + offset
        near_t = f32((y - f32(24.0))) / f32(dmin)  # walk t of the floor along that ray
        w.gi_update(p, dirs, bt, pal, pr)
        if near_t < f32(1.2):
            m = min(f32(1.2) - near_t, f32(0.35))
            expect = f32(np.clip(f32(off + f32(1.0) * m), f32(-3.5), f32(3.5)))
        elif near_t > f32(2.5):
            expect = f32(off * f32(0.98))
        else:
            expect = off
        got = f32(pr[0, 46])  # st_info.z = offset y
        assert abs(float(got) - float(expect)) < 2e-6, (step, got, expect)
        assert pr[0, 45] == 0.0 and pr[0, 47] == 0.0
        off = got


def test_fold_staging_by_hand():
    """epochs = 2: a new slot shows its first subset; the next round stages the other subset
    and, at the last epoch, folds the staged average with hysteresis 0.6."""
    w = oracle.OracleWorld((64, 64, 64))
    for x in range(64):
        for z in range(64):
            for y in range(20):
                w.set_voxel(x, y, z, 2 if y == 19 else 4)
    p = gi.gi_params(SUN35, grid=(1, 1, 1), spacing=64, rays_per_probe=32, epochs=2,
                     update_div=1, bounces=1)
    dirs, bt, pal = _tables(p)
    pr = _probes(p)
    origin = np.float32([32.0, 32.0, 32.0])

    def subset(epoch):
        c = np.zeros(20)
        for k in range(32):
            d = dirs[k * 2 + epoch, :3]
            rad, _, _, _ = w.gi_sample(p, bt, pal, origin, d, 0, k)
            d = d.astype(np.float64)
            b = np.array([0.282095, 0.488603 * d[1], 0.488603 * d[2], 0.488603 * d[0]])
            ws = [rad[0], rad[1], rad[2], rad[3], rad[3] * rad[3]]
            for a in range(5):
                c[4 * a:4 * a + 4] += (4 * np.pi / 32) * ws[a] * b
        return c

    s0, s1 = subset(0), subset(1)
    p["round"] = 0
    w.gi_update(p, dirs, bt, pal, pr)
    np.testing.assert_allclose(pr[0, :20], s0, rtol=1e-12)
    assert pr[0, 44] == 0
    p["round"] = 1  # epoch 1 = the last: stage s1, fold (s1 / 1) into the display
    w.gi_update(p, dirs, bt, pal, pr)
    h = float(np.float32(0.6))  # PROBE_FOLD_HYSTERESIS as stored (f32)
    np.testing.assert_allclose(pr[0, :20], s1 * (1 - h) + s0 * h, rtol=1e-12)
    assert np.all(pr[0, 24:44] == 0) and pr[0, 44] == 0
    p["round"] = 2  # epoch 0 again: staged, not folded
    w.gi_update(p, dirs, bt, pal, pr)
    np.testing.assert_allclose(pr[0, 24:44], s0, rtol=1e-12)
    assert pr[0, 44] == 1.0


def test_reference_estimator_equals_update_on_same_samples():
    """orc_gi_reference with the update's own direction set and round reproduces a new
    slot's first estimate exactly (the same samples, the same sums)."""
    w = oracle.OracleWorld((128, 128, 128)).generate(0x5EED_0001)
    w.rebuild()
    p = gi.gi_params(SUN35, grid=(4, 4, 4), spacing=32, rays_per_probe=64, epochs=1,
                     update_div=1, bounces=2)
    dirs, bt, pal = _tables(p)
    pr = _probes(p)
    slots = np.arange(64)
    probes0 = np.zeros(64, PROBE_DTYPE)
    origins = gi.probe_origins(probes0, p, slots)
    ref = w.gi_reference(p, dirs, bt, pal, slots, origins)
    w.gi_update(p, dirs, bt, pal, pr)
    live = pr[:, 23] == 1.0
    assert live.sum() > 8
    np.testing.assert_allclose(pr[live, :20], ref[live], rtol=1e-12, atol=1e-12)
    # buried slots are left empty with age 0
    assert np.all(pr[~live, :20] == 0) and np.all(pr[~live, 23] == 0)


def test_irradiance_error_metric():
    sh = np.zeros((2, 20))
    sh[:, 0] = sh[:, 4] = sh[:, 8] = 1.0 / 0.282095  # E/pi = 1 in every direction
    np.testing.assert_allclose(gi.sh_irradiance(sh), 1.0, rtol=1e-12)
    assert np.all(gi.irradiance_error(sh, sh) == 0)
    np.testing.assert_allclose(gi.irradiance_error(sh * 1.02, sh), 0.02, rtol=1e-9)
