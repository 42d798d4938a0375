"""vg_trace_frame_noalloc (the frame for callers that cannot allocate: its scratch is a static
per-device array) against vg_trace_frame, through the C ABI: the same bytes on two calls in a
row, the second with a smaller frame, then the size limit."""
import ctypes as C

import numpy as np
import pytest

import voxelg_amd as vg
from voxelg_amd import workloads
from voxelg_amd._lib import VGError, check, lib
from voxelg_amd.device import decode_hits

pytestmark = pytest.mark.gpu


def f32x3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def noalloc_frame(dw, cam, W, H, ptmax, flags, sun, stmax, hits, occ, stream):
    cam = np.ascontiguousarray(cam, np.float32)
    return lib().vg_trace_frame_noalloc(dw._h, cam.ctypes.data, W, H, ptmax, flags, f32x3(sun),
                                        stmax, hits.data_ptr(), occ.data_ptr(),
                                        stream.handle if stream is not None else None)


def test_noalloc_frame_matches_trace_frame(ctx, small):
    sun = workloads.LARGE.sun()
    stream = ctx.stream()
    cases = [((1280, 720), (128.5, 90.0, 128.5), 0.7, vg.VG_RAY_LOD | vg.VG_RAY_BEAM, 700.0),
             ((640, 360), (40.0, 75.0, 200.0), 2.4, vg.VG_RAY_BEAM, 512.0)]
    # outputs for both calls allocated up front: the calls themselves allocate nothing
    bufs = [(ctx.alloc(W * H * 20), ctx.alloc(W * H)) for (W, H), *_ in cases]
    for ((W, H), pos, yaw, flags, ptmax), (hits, occ) in zip(cases, bufs):
        print(f"case: noalloc frame {W}x{H} flags {flags}", flush=True)
        cam = vg.Camera(pos=pos, yaw=yaw, pitch=-0.35).uniform(W, H)
        hits.zero()
        occ.zero()
        ctx.synchronize()
        check(noalloc_frame(small, cam, W, H, ptmax, flags, sun, 480.0, hits, occ, stream),
              "vg_trace_frame_noalloc")
        stream.synchronize()
        ref_h, ref_o = small.trace_frame(cam, W, H, ptmax, flags, sun, 480.0)
        ctx.synchronize()
        got = decode_hits(hits)
        assert got.tobytes() == decode_hits(ref_h).tobytes(), f"{W}x{H}: hits differ"
        assert np.array_equal(occ.numpy(np.uint8), ref_o.numpy(np.uint8)), f"{W}x{H}: shadows"
        assert got["hit"].mean() > 0.2


def test_noalloc_frame_rejects_frames_above_its_scratch(ctx, small):
    W, H = 4104, 4096  # 513 x 512 blocks of 8x8 pixels: one column above the limit
    cam = vg.Camera(pos=(128.5, 90.0, 128.5), yaw=0.7, pitch=-0.35).uniform(W, H)
    st = lib().vg_trace_frame_noalloc(small._h, np.ascontiguousarray(cam).ctypes.data, W, H,
                                      512.0, 0, f32x3(workloads.LARGE.sun()), 480.0, None, None,
                                      None)
    assert st != 0
    with pytest.raises(VGError) as e:
        check(st, "vg_trace_frame_noalloc")
    assert "scratch" in str(e.value), str(e.value)  # refused for its size, before any launch
