"""GPU parity of the water secondary rays (vg_trace_water_secondary): for every water pixel
of a traced frame, the reflection and refraction hit records equal the oracle's R1 traces of
the same rays (orc_water_rays) bit for bit; other pixels get misses."""
import math

import numpy as np
import pytest

import oracle
import voxelg_amd as vg
from voxelg_amd._lib import VGError
from voxelg_amd.device import decode_hits

pytestmark = pytest.mark.gpu

SUN = vg.sun_direction(math.radians(35.0), 2.1)


@pytest.mark.parametrize("pose", [((228.28, 80.0, 236.28), 3.927, -0.5),
                                  ((128.5, 90.0, 128.5), 0.7, -0.35)])
def test_water_secondary_matches_oracle(ctx, small, small_world, pose):
    W, H = 192, 108
    cam = vg.Camera(pos=pose[0], yaw=pose[1], pitch=pose[2]).uniform(W, H)
    print(f"case: water secondary rays, camera {pose}", flush=True)
    hits, _ = small.trace_frame(cam, W, H, 512.0, 0, SUN, 480.0)
    refl, refr = small.trace_water_secondary(cam, W, H, hits)
    ctx.synchronize()
    prim = decode_hits(hits)
    ow = oracle.OracleWorld.view(small_world)
    erefl, erefr = ow.water_secondary(cam, W, H, prim)
    got_refl, got_refr = decode_hits(refl), decode_hits(refr)
    assert got_refl.tobytes() == erefl.tobytes()
    assert got_refr.tobytes() == erefr.tobytes()
    water = (prim["hit"] == 1) & (prim["material"] >= 5) & (prim["material"] <= 12)
    if pose[1] > 3.0:
        assert water.sum() > 1000 and got_refr["hit"][water].mean() > 0.5


def test_water_secondary_errors(ctx, small):
    cam = vg.Camera(pos=(128.5, 90.0, 128.5), yaw=0.7, pitch=-0.35).uniform(64, 32)
    hits, _ = small.trace_frame(cam, 64, 32, 512.0, 0, SUN, 480.0)
    with pytest.raises(VGError):
        small.trace_water_secondary(cam, 32, 32, hits)  # resolution mismatch
