"""Pins of the oracle's probe-cache consumer (sample_probes / sh_eval_irradiance,
shaders/gi_probes.wgsl:69-179; indirect_light in probe mode, shaders/rt_gi.wgsl:128-139)
against closed forms: SH-L1 irradiance with and without deringing, a constant probe field,
two-probe blends at geometry where the smoothstepped trilinear weights are exactly 0, 1/2 or
1 (so the backface and Chebyshev factors reduce to their constants), ungathered probes, and
the distance fade."""
import numpy as np
import pytest

import oracle
from voxelg_amd._lib import PROBE_DTYPE

GRID, SPACING = (4, 4, 4), 8
A1 = np.float32(0.488603 * 0.666667)
C0 = np.float32(0.282095)


def empty_probes():
    return np.zeros(GRID[0] * GRID[1] * GRID[2], PROBE_DTYPE)


def gather(probes, g, dc=(1.0, 1.0, 1.0), dist_dc=100.0, dist2_dc=0.0):
    """Mark probe cell g gathered (info = g, age 1) with a DC-only radiance and distance
    moments."""
    i = g[0] + g[1] * GRID[0] + g[2] * GRID[0] * GRID[1]
    for c, name in enumerate(("sh_r", "sh_g", "sh_b")):
        probes[i][name] = (dc[c], 0.0, 0.0, 0.0)
    probes[i]["sh_dist"] = (dist_dc, 0.0, 0.0, 0.0)
    probes[i]["sh_dist2"] = (dist2_dc, 0.0, 0.0, 0.0)
    probes[i]["info"] = (g[0], g[1], g[2], 1.0)
    return probes


def rec(sh_r, sh_g=(0, 0, 0, 0), sh_b=(0, 0, 0, 0)):
    p = np.zeros(1, PROBE_DTYPE)
    p[0]["sh_r"], p[0]["sh_g"], p[0]["sh_b"] = sh_r, sh_g, sh_b
    return p


def test_sh_irradiance_dc_and_deringing():
    # DC only: E = c * 0.282095 in every direction
    for n in ((1, 0, 0), (0, -1, 0), (0.6, 0.0, 0.8)):
        e = oracle.sh_irradiance(rec((2.0, 0, 0, 0)), n)
        assert e[0] == pytest.approx(2.0 * C0, rel=1e-7) and e[1] == 0 and e[2] == 0
    # a weak lobe along +y (stored c1y first) evaluates as plain SH: dc + a1 * (c1 . n)
    e = oracle.sh_irradiance(rec((4.0, 0.5, 0, 0)), (0, 1, 0))
    assert e[0] == pytest.approx(4.0 * C0 + 0.5 * A1, rel=1e-6)
    # a strong lobe is capped at 95 % of the DC term: 1.95 dc along it, 0.05 dc against it,
    # dc across it
    dc = 1.0 * C0
    strong = rec((1.0, 5.0, 0, 0))
    assert oracle.sh_irradiance(strong, (0, 1, 0))[0] == pytest.approx(1.95 * dc, rel=1e-6)
    assert oracle.sh_irradiance(strong, (0, -1, 0))[0] == pytest.approx(0.05 * dc, rel=1e-5)
    assert oracle.sh_irradiance(strong, (1, 0, 0))[0] == pytest.approx(dc, rel=1e-6)
    # stored order (c1y, c1z, c1x): a +x lobe sits in the last coefficient
    assert oracle.sh_irradiance(rec((1.0, 0, 0, 5.0)), (1, 0, 0))[0] == pytest.approx(1.95 * dc, rel=1e-6)
    # a negative field clamps to zero
    assert oracle.sh_irradiance(rec((-1.0, 0, 0, 0)), (0, 1, 0))[0] == 0.0


def test_constant_field_returns_the_field():
    probes = empty_probes()
    for x in range(4):
        for y in range(4):
            for z in range(4):
                gather(probes, (x, y, z), dc=(1.0, 2.0, 3.0))
    rng = np.random.default_rng(3)
    for _ in range(50):
        pw = rng.uniform(6.0, 26.0, 3)
        n = np.zeros(3)
        n[rng.integers(3)] = rng.choice((-1.0, 1.0))
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        e = oracle.sample_probes(probes, GRID, SPACING, pw, n, v)
        np.testing.assert_allclose(e, np.array((1.0, 2.0, 3.0)) * C0, rtol=2e-6)


def test_ungathered_and_foreign_probes_are_skipped():
    probes = empty_probes()
    e = oracle.sample_probes(probes, GRID, SPACING, (16.0, 14.0, 16.0), (0, 1, 0), (0, -1, 0))
    assert np.all(e == 0)  # nothing gathered: zero
    # a slot holding a different world cell (info mismatch) counts as not gathered
    gather(probes, (1, 1, 1))
    probes[1 + 1 * 4 + 1 * 16]["info"] = (5.0, 1.0, 1.0, 1.0)
    e = oracle.sample_probes(probes, GRID, SPACING, (12.0, 10.0, 12.0), (0, 1, 0), (0, -1, 0))
    assert np.all(e == 0)


def two_probe_setup(dist_dc_above=100.0, dist2_dc_above=0.0):
    """Probe A below the biased query point, probe B above it, both on its x/z lattice
    lines: the query point p_q = pw + n - v = (12, 16, 12) sits half way (y weight 1/2 each,
    x and z weights exactly 1)."""
    probes = empty_probes()
    gather(probes, (1, 1, 1), dc=(1.0, 1.0, 1.0))                      # A at (12, 12, 12)
    gather(probes, (1, 2, 1), dc=(3.0, 3.0, 3.0), dist_dc=dist_dc_above,
           dist2_dc=dist2_dc_above)                                    # B at (12, 20, 12)
    return probes, (12.0, 14.0, 12.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0)


def test_backface_weights():
    probes, pw, n, v = two_probe_setup()
    e = oracle.sample_probes(probes, GRID, SPACING, pw, n, v)
    # A lies behind the surface (backface 0: weight 1/2 * 0.05), B in front (1/2 * 1.05);
    # both far inside their mean visible distance (no Chebyshev factor)
    wa, wb = 0.5 * 0.05, 0.5 * 1.05
    expect = (1.0 * C0 * wa + 3.0 * C0 * wb) / (wa + wb)
    np.testing.assert_allclose(e, [expect] * 3, rtol=2e-6)


def test_chebyshev_visibility():
    # B's mean visible distance toward the point is 1 (mu = 1) with mu2 = 1: the point lies
    # 4 away, delta = 4 - (1 + 0.35) = 2.65, variance floored at 0.25
    probes, pw, n, v = two_probe_setup(dist_dc_above=1.0 / float(C0), dist2_dc_above=1.0 / float(C0))
    e = oracle.sample_probes(probes, GRID, SPACING, pw, n, v)
    cheb = 0.25 / (0.25 + 2.65 ** 2)
    wa, wb = 0.5 * 0.05, 0.5 * 1.05 * cheb
    expect = (1.0 * C0 * wa + 3.0 * C0 * wb) / (wa + wb)
    np.testing.assert_allclose(e, [expect] * 3, rtol=2e-5)


def test_frame_fade_and_misses():
    # a constant field 256 voxels wide, the camera in its middle looking along +x: hit
    # points up to t = 90 stay inside it, so a pixel's value is C0 times the fade
    grid, spacing = (32, 4, 32), 8
    probes = np.zeros(grid[0] * grid[1] * grid[2], PROBE_DTYPE)
    for z in range(grid[2]):
        for y in range(grid[1]):
            for x in range(grid[0]):
                i = x + y * grid[0] + z * grid[0] * grid[1]
                for name in ("sh_r", "sh_g", "sh_b"):
                    probes[i][name] = (1.0, 0.0, 0.0, 0.0)
                probes[i]["sh_dist"] = (1000.0, 0.0, 0.0, 0.0)
                probes[i]["info"] = (x, y, z, 1.0)
    from voxelg_amd import Camera
    W, H = 6, 1
    cam = Camera(pos=(128.0, 16.0, 128.0), yaw=0.0, pitch=0.0).uniform(W, H)
    hits = np.zeros(W * H, np.dtype([("voxel", "<i4", (3,)), ("t", "<f4"), ("face", "u1"),
                                      ("material", "u1"), ("hit", "u1"), ("pad", "u1")]))
    hits["t"] = (1.0, 1.0, 54.0, 72.0, 90.0, 200.0)
    hits["face"] = 2
    hits["hit"] = (1, 0, 1, 1, 1, 1)
    out = oracle.frame_indirect(probes, grid, spacing, cam, W, H, hits)
    assert np.all(out[:, 3] == 0)
    assert np.all(out[1] == 0)  # a miss
    # 1 - smoothstep(54, 90, t): 1 up to 54, 1/2 at 72 (smoothstep's midpoint), 0 from 90 on
    for px, fade in ((0, 1.0), (2, 1.0), (3, 0.5), (4, 0.0), (5, 0.0)):
        np.testing.assert_allclose(out[px, :3], [float(C0) * fade] * 3, rtol=2e-6, atol=0)
