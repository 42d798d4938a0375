"""The reference's known-answer scenes (src/accel.rs:515-942), rebuilt on both worlds.

Each builder returns an oracle.OracleWorld (reference-semantics set_voxel / apply_edit with
incremental masks) plus a voxelg_amd.World holding the same bytes for upload.
"""
import numpy as np

import oracle
import voxelg_amd as vg

REF_DIMS = (512, 256, 512)  # World::new() window, src/world_dims.rs:13-15
STONE = 4


def to_world(ow: oracle.OracleWorld) -> vg.World:
    w = vg.World(ow.dims, ow.origin)
    w.bricks[:] = ow.bricks
    w.tile_mask[:] = ow.tile_mask
    w.chunk_mask[:] = ow.chunk_mask
    w.l4_mask[:] = ow.l4_mask[: len(w.l4_mask)]
    w.brick_uniform[:] = ow.brick_uniform
    w.tile_uniform[:] = ow.tile_uniform
    return w


def fill_box(ow, x0, x1, y0, y1, z0, z1, mat, world_coords=True):
    """Set every voxel of [x0,x1) x [y0,y1) x [z0,z1) (world coords) to mat."""
    for z in range(z0, z1):
        for y in range(y0, y1):
            for x in range(x0, x1):
                ow.apply_edit(x, y, z, mat)


def flat_floor(x0, z0, n, origin=(0, 0, 0)):
    """src/accel.rs:515-526: stone slab, top solid layer y = 60, x,z in [x0, x0+n)."""
    ow = oracle.OracleWorld(REF_DIMS, origin)
    fill_box(ow, x0 + origin[0], x0 + n + origin[0], 57, 61, z0 + origin[2], z0 + n + origin[2],
             STONE)
    return ow


PILLARS = [(20, 24, 90), (40, 40, 100), (52, 18, 76), (30, 50, 84)]


def floor_pillars(origin=(0, 0, 0)):
    """src/accel.rs:601-621: floor at local y 57..60, x,z in [8,64), four 3x3 pillars."""
    ox, oy, oz = origin
    ow = oracle.OracleWorld(REF_DIMS, origin)
    fill_box(ow, ox + 8, ox + 64, oy + 57, oy + 61, oz + 8, oz + 64, STONE)
    for px, pz, top in PILLARS:
        fill_box(ow, ox + px, ox + px + 3, oy + 61, oy + top, oz + pz, oz + pz + 3, STONE)
    return ow


def floor_pillar_rays(origin=(0, 0, 0)):
    """src/accel.rs:623-653: 4 high eyes x 8x8 floor targets + 4 pillars x 3x3 side targets,
    as (world eye, unit dir) pairs."""
    base = np.asarray(origin, np.float32)
    pairs = []
    for eye in [(6.0, 96.0, 6.0), (70.0, 110.0, 12.0), (34.0, 130.0, 34.0), (12.0, 74.0, 60.0)]:
        for gz in range(8):
            for gx in range(8):
                tgt = np.array([12.0 + gx * 6.0, 62.0, 12.0 + gz * 6.0], np.float32)
                pairs.append((np.asarray(eye, np.float32), tgt))
    for px, pz, top in PILLARS:
        eye = np.array([px - 16.0, (61 + top) * 0.5, pz + 1.5], np.float32)
        for gy in range(3):
            for gz in range(3):
                tgt = np.array([px, 61.0 + (top - 61) * (0.25 + gy * 0.25), pz + 0.5 + gz],
                               np.float32)
                pairs.append((eye, tgt))
    eyes, dirs = [], []
    for eye, tgt in pairs:
        d = tgt - eye
        d = (d * np.float32(1.0 / np.sqrt(np.float32(np.dot(d, d))))).astype(np.float32)
        eyes.append(eye + base)
        dirs.append(d)
    return np.array(eyes, np.float32), np.array(dirs, np.float32)


def lattice_touch(p, eps=0.02):
    """src/accel.rs:588-599: how many coordinates of p sit on a voxel-lattice plane."""
    p = np.asarray(p, np.float64)
    return int(np.sum(np.abs(p - np.round(p)) < eps))


def ray_box_entry(o, d, v):
    """src/accel.rs:578-586 in fp32."""
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    with np.errstate(divide="ignore"):
        inv = np.where(np.abs(d) < 1e-8, np.float32(1e30), np.float32(1.0) / d).astype(np.float32)
    vmin = np.asarray(v, np.float32)
    t0 = (vmin - o) * inv
    t1 = (vmin + np.float32(1.0) - o) * inv
    return max(float(np.max(np.minimum(t0, t1))), 0.0)


def classify(eyes, dirs, cand_hits, cpu_hits):
    """The accel_matches_cpu_raycast classification (src/accel.rs:655-718) of candidate hits
    against CPU raycast hits. Returns (clean, floor, pillar, grazing, disagreements)."""
    clean = floor = pillar = grazing = 0
    bad = []
    for i in range(len(eyes)):
        h, c = cand_hits[i], cpu_hits[i]
        if not h["hit"] and not c["hit"]:
            continue
        cpu_p = eyes[i] + ray_box_entry(eyes[i], dirs[i], c["voxel"]) * dirs[i] if c["hit"] else None
        cand_p = eyes[i] + np.float32(h["t"]) * dirs[i] if h["hit"] else None
        cpu_graze = cpu_p is not None and lattice_touch(cpu_p) >= 2
        cand_graze = cand_p is not None and lattice_touch(cand_p) >= 2
        agree = h["hit"] and c["hit"] and tuple(h["voxel"]) == tuple(c["voxel"])
        if agree:
            if cpu_graze:
                grazing += 1
            else:
                clean += 1
                if h["face"] != c["face"]:
                    bad.append((i, "face", h, c))
                if h["voxel"][1] <= 60:
                    floor += 1
                else:
                    pillar += 1
            continue
        if cpu_graze or cand_graze:
            grazing += 1
            continue
        bad.append((i, "voxel", h, c))
    return clean, floor, pillar, grazing, bad


def roof_scene():
    """src/accel.rs:836-851: flat floor [8,64) + stone roof slab y = 80 over [24,44)^2."""
    ow = flat_floor(8, 8, 56)
    fill_box(ow, 24, 44, 80, 81, 24, 44, STONE)
    return ow


def roof_samples():
    """src/accel.rs:853-874: floor-cell centres just above the top face, classified by where
    the sun ray crosses the roof plane (margin 1.5)."""
    sun = np.array([0.16, 0.95, 0.11], np.float32)
    sun = (sun * np.float32(1.0 / np.sqrt(np.float32(np.dot(sun, sun))))).astype(np.float32)
    origins, under, open_ = [], [], []
    for z in range(10, 62):
        for x in range(10, 62):
            o = np.array([x + 0.5, 61.02, z + 0.5], np.float32)
            s = (80.0 - o[1]) / sun[1]
            hx, hz = o[0] + sun[0] * s, o[2] + sun[2] * s
            m = 1.5
            origins.append(o)
            under.append(24 + m < hx < 44 - m and 24 + m < hz < 44 - m)
            open_.append(hx < 24 - m or hx > 44 + m or hz < 24 - m or hz > 44 + m)
    return np.array(origins, np.float32), sun, np.array(under), np.array(open_)


def assert_agrees_with_r0(c, what):
    """A classify_vs_r0 result meets the R0 agreement criterion: hits match R0 on voxel, face
    and material, t within 1e-4 voxels, and the only disagreements are edge ties
    (lattice_touch >= 2) or tmax-boundary ties, rare ones."""
    assert c["clean_disagree"] == 0, (what, c)
    assert c["face_mismatch"] == 0 and c["material_mismatch"] == 0, (what, c)
    assert c["max_dt"] <= 1e-4, (what, c)
    assert c["disagree_ties"] + c["tmax_ties"] <= 1e-4 * c["rays"] + 16, (what, c)
    print(what, {k: v for k, v in c.items() if k != "bad_index"}, flush=True)


def same_shadow_ray(cand, ref, eye, dirs, sel):
    """Rays of sel whose primary hits agree with R0's on voxel, face and entry t: both sides
    then trace the same shadow ray (an edge tie entered through another face would move the
    shadow origin)."""
    return sel & (cand["hit"] != 0) & (ref["hit"] != 0) & \
        np.all(cand["voxel"] == ref["voxel"], axis=1) & (cand["face"] == ref["face"]) & \
        (cand["t"] == ref["t"])
