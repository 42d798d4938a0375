"""Pin the oracle's ray generators by geometry, not by their own formulas:
  - orc_primary_dir (pinhole rays, shaders/raymarch.wgsl:430-438, 600; src/camera.rs:25-113):
    square pixels on the image plane, the horizontal and vertical fields of view, which way is
    up, left-right mirror symmetry, the jitter as a pixel shift, aspect cases W > H and W < H;
  - orc_shadow_origin (p_hit + n * 0.001, shaders/raymarch.wgsl:3955): every origin lies
    0.001 outside the hit face, inside the empty voxel the ray came through, on the face.
The GPU reproduces both bit for bit (frame parity tests), so these pin the device path too."""
import math

import numpy as np
import pytest

import oracle
import voxelg_amd as vg
from voxelg_amd import workloads


def components(cam_u, dirs):
    """(along forward, along right, along up) of unit directions, in float64."""
    d = np.asarray(dirs, np.float64).reshape(-1, 3)
    f, r, u = (np.asarray(cam_u[a:a + 3], np.float64) for a in (4, 8, 12))
    return d @ f, d @ r, d @ u


@pytest.mark.parametrize("wh", [(33, 19), (19, 33), (64, 64), (128, 37)])
@pytest.mark.parametrize("fov_deg", [70.0, 30.0])
def test_pinhole_geometry(wh, fov_deg):
    W, H = wh
    cam = vg.Camera(pos=(3.0, 4.0, 5.0), yaw=0.9, pitch=-0.3, fov_y=math.radians(fov_deg))
    u = cam.uniform(W, H)
    tan_half = math.tan(math.radians(fov_deg) / 2)
    d = oracle.primary_dirs(u, W, H).reshape(H, W, 3)
    assert np.allclose(np.linalg.norm(d, axis=2), 1.0, atol=2e-7)
    df, dr, du = (c.reshape(H, W) for c in components(u, d))
    assert (df > 0).all()
    # image-plane coordinates at unit distance along forward
    x, y = dr / df, du / df
    # square pixels: one column to the right and one row down move the same distance on the
    # image plane, 2 tan(fov_y / 2) / H (a dropped aspect would make the column step / W)
    step = 2 * tan_half / H
    assert np.allclose(np.diff(x, axis=1), step, atol=2e-6)
    assert np.allclose(np.diff(y, axis=0), -step, atol=2e-6)
    assert np.allclose(np.diff(x, axis=0), 0.0, atol=2e-6)
    assert np.allclose(np.diff(y, axis=1), 0.0, atol=2e-6)
    # the first row looks up, the last down (a flipped ndc.y would swap them)
    assert (du[0] > 0).all() and (du[-1] < 0).all()
    # the first column looks left of forward, the last right
    assert (dr[:, 0] < 0).all() and (dr[:, -1] > 0).all()
    # fields of view: the top row's centre ray sits half a pixel inside fov_y / 2, the left
    # column's half a pixel inside the horizontal half-angle atan(tan(fov_y / 2) W / H)
    top = math.atan(tan_half * (1 - 1 / H))
    assert np.allclose(np.arctan(y[0]), top, atol=2e-6)
    left = math.atan(tan_half * (W / H) * (1 - 1 / W))
    assert np.allclose(np.arctan(-x[:, 0]), left, atol=2e-6)
    # mirror symmetry about the forward-up plane
    assert np.allclose(dr, -dr[:, ::-1], atol=2e-6)
    assert np.allclose(du, du[:, ::-1], atol=2e-6) and np.allclose(df, df[:, ::-1], atol=2e-6)
    if W % 2 and H % 2:
        assert np.allclose(d[H // 2, W // 2], u[4:7], atol=1e-6)


def test_pinhole_jitter_is_a_pixel_shift():
    W, H = 40, 24
    u = vg.Camera(pos=(0, 0, 0), yaw=2.2, pitch=0.4).uniform(W, H)
    d = oracle.primary_dirs(u, W, H).reshape(H, W, 3)
    uj = u.copy()
    uj[24], uj[25] = 1.0, 1.0  # jitter of one pixel right and one pixel down
    dj = oracle.primary_dirs(uj, W, H).reshape(H, W, 3)
    assert dj[:-1, :-1].tobytes() == d[1:, 1:].tobytes()
    uh = u.copy()
    uh[24] = 0.5  # half a pixel: halfway on the image plane
    dh = oracle.primary_dirs(uh, W, H).reshape(H, W, 3)
    fh, rh, _ = components(u, dh)
    f0, r0, _ = components(u, d)
    x0, xh = (r0 / f0).reshape(H, W), (rh / fh).reshape(H, W)
    assert np.allclose(xh[:, :-1], 0.5 * (x0[:, :-1] + x0[:, 1:]), atol=2e-6)


def test_shadow_origins_leave_the_hit_face(small_world):
    wl = workloads.SMALL
    _, cam = wl.camera()
    ow = oracle.OracleWorld.view(small_world)
    rows = (200, 260)
    hits, _, _ = ow.frame(cam, wl.width, wl.height, wl.primary_tmax, 0, None, 0.0, rows)
    dirs = oracle.primary_dirs(cam, wl.width, wl.height)[rows[0] * wl.width:rows[1] * wl.width]
    sel = np.flatnonzero(hits["hit"] != 0)[::7]
    h = hits[sel]
    so = oracle.shadow_origins(cam[:3], dirs[sel], h).astype(np.float64)
    face = h["face"].astype(int)
    assert (face < 6).all()
    ax = face >> 1
    n = np.where(face & 1, -1, 1)  # +x 0, -x 1, +y 2, -y 3, +z 4, -z 5
    rows_i = np.arange(len(sel))
    v = h["voxel"].astype(np.int64)
    plane = v[rows_i, ax] + (n > 0)
    # 0.001 outside the face plane along the outward normal (fp32 at these coordinates: the
    # offset is exact to ~3e-5)
    assert np.allclose((so[rows_i, ax] - plane) * n, 0.001, atol=1e-4)
    # hence in the neighbour voxel the ray came through, not in the hit voxel
    assert (np.floor(so[rows_i, ax]).astype(np.int64) == v[rows_i, ax] + n).all()
    # and on the face: within the voxel's extent on the other two axes
    for k in range(3):
        on = ax != k
        assert (so[on, k] >= v[on, k] - 1e-4).all() and (so[on, k] <= v[on, k] + 1 + 1e-4).all()
    # that neighbour is empty: a first hit is entered from empty space
    nb = v.copy()
    nb[rows_i, ax] += n
    mats = [ow.material_at(*p) for p in nb[:2000]]
    assert not any(m not in (0, 33) for m in mats)
    # top faces and both faces of at least one horizontal axis occur in the frame
    fs = set(int(f) for f in np.unique(face))
    assert 2 in fs and ({0, 1} <= fs or {4, 5} <= fs), fs
