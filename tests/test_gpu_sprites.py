"""GPU parity of the depth-sprite pass (DepthRenderPass, depth.vert/frag):
the sprite depth image equals the oracle's bit for bit (the per-pixel minimum
does not depend on the order of the particles; the GPU's early depth test
skips only atomics that cannot lower it), for the reference's perspective and
an oblique projection, and the reference march
started from the GPU's sprite depths (PerPixel_Isotropic's depth-buffer start,
RayMarcher.cpp:256-344) equals the oracle's march started from its own."""
import time

import numpy as np
import pytest

import oracle as O
from pmarch import scenes
from pmarch.camera import make_camera

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cam_i", [0, 1])
def test_depth_sprites_bitexact(marcher, cam_i):
    sc = scenes.lattice_cube(0)
    cam = sc.cameras[cam_i]
    print(f"depth sprites lattice_cube cam={cam_i}", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    got = marcher.depth_sprites(0, cam, sc.h)
    want = O.depth_sprites(sc.xyz, cam, sc.h)
    assert np.array_equal(got, want)
    assert 0.2 < (want < 1.0).mean() < 0.95


def test_depth_sprites_oblique_projection(marcher):
    """A projection whose depth also depends on x: the pass takes any View and
    Projection, not only the reference's perspective."""
    import dataclasses
    sc = scenes.lattice_cube(0)
    cam = sc.cameras[0]
    proj = cam.proj.copy()
    proj[2] = np.float32(0.02)                         # clip z += 0.02 x_view
    ob = dataclasses.replace(cam, proj=proj)
    print("depth sprites oblique projection", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    got = marcher.depth_sprites(0, ob, sc.h)
    assert np.array_equal(got, O.depth_sprites(sc.xyz, ob, sc.h))
    assert not np.array_equal(got, marcher.depth_sprites(0, cam, sc.h))


def test_depth_sprites_dam_and_edges(marcher):
    """The dam break at 480x270, and a camera inside the particle cloud (quads
    behind the eye, huge quads near it)."""
    sc = scenes.dam_break(0, nx=40, ny=80, nz=40, width=480, height=270)
    print("depth sprites dam break 480x270 + inside camera", flush=True)
    marcher.set_frame(1, sc.xyz, sc.h)
    for cam in (sc.cameras[0], make_camera(sc.xyz.mean(0), sc.xyz.mean(0) + [0.0, 0.0, 1.0], 203, 77)):
        got = marcher.depth_sprites(1, cam, sc.h)
        assert np.array_equal(got, O.depth_sprites(sc.xyz, cam, sc.h))


def test_depth_sprites_large_footprints(marcher):
    """A camera inside the lattice cube at 640x360: most quads cover more than
    1024 pixels (about 1,800 on average), so most sprites are listed and drawn
    by a whole wave, the rest by their own thread; the image equals the
    oracle's.  Then the 1 M dam break from inside at 1920x1080 (timed only)."""
    sc = scenes.lattice_cube(0)
    c = sc.xyz.mean(0)
    cam = make_camera(c, c + np.array([0.3, 0.2, 1.0]), 640, 360)
    print("depth sprites: camera inside the lattice cube, 640x360", flush=True)
    marcher.set_frame(2, sc.xyz, sc.h)
    got = marcher.depth_sprites(2, cam, sc.h)
    assert np.array_equal(got, O.depth_sprites(sc.xyz, cam, sc.h))
    assert (got < 1.0).mean() > 0.9
    dam = scenes.dam_break(0)
    inside = make_camera(dam.xyz.mean(0), dam.xyz.mean(0) + np.array([0.0, 0.0, 1.0]), 1920, 1080)
    marcher.set_frame(2, dam.xyz, dam.h)
    marcher.depth_sprites(2, inside, dam.h)
    t0 = time.perf_counter()
    img = marcher.depth_sprites(2, inside, dam.h)
    print(f"1 M dam break, camera inside, 1920x1080: {1e3 * (time.perf_counter() - t0):.2f} ms "
          f"(host image out), {(img < 1.0).mean():.3f} of the pixels drawn", flush=True)
    assert (img < 1.0).mean() > 0.9


def test_reference_march_from_sprites(marcher):
    sc = scenes.lattice_cube(1)
    cam = sc.cameras[0]
    print("reference march from sprite depths", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    depth = marcher.depth_sprites(0, cam, sc.h)
    st = dict(mode=1, max_steps=128, step=0.009, iso=sc.iso, normal_mode=1, skip_mode=0, compensated=0)
    pos, nrm, t, stats = marcher.march(0, st, cam.inv_pv, cam.eye, cam.width, cam.height, depth=depth)
    G = O.Grid(sc.xyz, sc.h)
    opos, onrm, ot, ocnt = O.march(G, O.make_settings(**st), cam.inv_pv, cam.eye, cam.width, cam.height,
                                   depth=O.depth_sprites(sc.xyz, cam, sc.h))
    assert np.array_equal(pos, opos)
    assert np.array_equal(nrm, onrm, equal_nan=True)
    assert stats["hits"] == ocnt["hits"] and stats["hits"] > 10000


def test_sprites_to_march_on_device(marcher):
    """pm_depth_sprites_device then pm_march_async_device reading those depths:
    the reference pipeline with no host copy equals the host hand-off; a host
    address for the depth image is refused before any launch."""
    from pmarch import PmError
    from test_gpu_device_out import _alloc, _not_device, _to_host
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipPointerGetAttributes.argtypes = [C.c_void_p, C.c_void_p]
    sc = scenes.lattice_cube(2, width=203, height=77)
    cam = sc.cameras[0]
    W, H = cam.width, cam.height
    print("sprites -> march on device 203x77", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    st = dict(mode=1, max_steps=128, step=0.009, iso=sc.iso, normal_mode=1, skip_mode=0, compensated=0)
    depth = marcher.depth_sprites(0, cam, sc.h)
    want = marcher.march(0, st, cam.inv_pv, cam.eye, W, H, depth=depth)
    bufs = [_alloc(hip, 4 * W * H), _alloc(hip, 16 * W * H), _alloc(hip, 16 * W * H)]
    try:
        marcher.depth_sprites_device(0, cam, sc.h, bufs[0])
        marcher.march_device(0, st, cam.inv_pv, cam.eye, W, H, bufs[1], bufs[2], 0, d_depth=bufs[0])
        assert np.array_equal(_to_host(hip, bufs[0], (H, W)), depth)
        assert np.array_equal(_to_host(hip, bufs[1], (H, W, 4)), want[0])
        assert np.array_equal(_to_host(hip, bufs[2], (H, W, 4)), want[1], equal_nan=True)
        host = np.zeros((H, W), np.float32)
        assert _not_device(hip, host.ctypes.data)
        with pytest.raises(PmError):
            marcher.depth_sprites_device(0, cam, sc.h, host.ctypes.data)
    finally:
        for b in bufs:
            hip.hipFree(b)


def test_sprites_device_completion_and_rebuild(marcher):
    """pm_depth_sprites_device is complete once pm_wait returns: a null-stream
    copy afterwards reads the whole image (equal to the host pass).  A frame
    rebuilt on the build stream (pm_set_frame_stream) right after a sprite pass
    of that frame starts after the pass: the image is that of the old
    particles, and the new frame then draws its own."""
    from test_gpu_device_out import _alloc, _to_host
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    sc = scenes.dam_break(0, nx=40, ny=80, nz=40, width=480, height=270)
    moved = (sc.xyz + np.float32(0.05)).astype(np.float32)
    cam = sc.cameras[0]
    W, H = cam.width, cam.height
    print("sprites on device: pm_wait, then a rebuild of the same frame 480x270", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    want_old = marcher.depth_sprites(0, cam, sc.h)
    d = _alloc(hip, 4 * W * H)
    try:
        marcher.depth_sprites_device(0, cam, sc.h, d)
        marcher.wait()
        assert np.array_equal(_to_host(hip, d, (H, W)), want_old)
        marcher.depth_sprites_device(0, cam, sc.h, d)
        marcher.set_frame_stream(0, moved, sc.h)
        marcher.wait()
        assert np.array_equal(_to_host(hip, d, (H, W)), want_old)
        want_new = marcher.depth_sprites(0, cam, sc.h)
        assert not np.array_equal(want_new, want_old)
        assert np.array_equal(want_new, O.depth_sprites(moved, cam, sc.h))
    finally:
        hip.hipFree(d)
