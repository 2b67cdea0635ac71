"""pm_march_device_noalloc -- no counterpart in the reference; the march for
callers that cannot allocate device memory: the images it writes into the
caller's device memory, with the library's static scratch, equal
pm_march_async_device's bit for bit, on two calls in a row (the second a
smaller image, so the scratch holds the first call's records), for the plain
and the learned march; a size beyond the scratch, and bound bits not built yet
(for the march's iso and bound_boxes), are refused before anything is queued."""
import ctypes as C

import numpy as np
import pytest

from pmarch import scenes, train
from pmarch.camera import make_camera, orbit_eye

pytestmark = pytest.mark.gpu

PM_ERR_SIZE, PM_ERR_STATE = -4, -6


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return lib


def _alloc(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p.value


def _to_host(hip, ptr, shape):
    out = np.zeros(shape, np.float32)
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0        # device to host
    return out


def _images(hip, bufs, W, H):
    return (_to_host(hip, bufs[0], (H, W, 4)), _to_host(hip, bufs[1], (H, W, 4)), _to_host(hip, bufs[2], (H, W)))


@pytest.mark.parametrize("learned", [0, 1])
def test_noalloc_matches_device_march(marcher, hip, learned):
    sc = scenes.dam_break(2, nx=40, ny=80, nz=40, width=640, height=360)
    target = (0.2, 0.4, 0.2)
    marcher.set_frame(60, sc.xyz, sc.h)
    st = sc.settings(skip_mode=4, learned=learned)
    if learned:
        params, _ = train.load()
        marcher.mlp_set(params)
    # the bound bits for this iso, built by an ordinary call
    marcher.bound_grid(60, sc.iso)
    for W, H in ((640, 360), (203, 77)):
        cam = make_camera(orbit_eye(target, 1.1, 25.0, 40.0), target, W, H)
        print(f"noalloc learned={learned} {W}x{H}", flush=True)
        ref = [_alloc(hip, 16 * W * H), _alloc(hip, 16 * W * H), _alloc(hip, 4 * W * H)]
        got = [_alloc(hip, 16 * W * H), _alloc(hip, 16 * W * H), _alloc(hip, 4 * W * H)]
        try:
            s_ref = marcher.march_device(60, st, cam.inv_pv, cam.eye, W, H, ref[0], ref[1], ref[2])
            s_got = marcher.march_device_noalloc(60, st, cam.inv_pv, cam.eye, W, H, got[0], got[1], got[2])
            a, b = _images(hip, ref, W, H), _images(hip, got, W, H)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
            assert s_ref["hits"] > 100
            for k in ("rays", "rays_in_aabb", "hits", "samples", "march_samples", "march_cand", "cert_skipped"):
                assert s_ref[k] == s_got[k], k
        finally:
            for p in ref + got:
                hip.hipFree(p)


def test_noalloc_refuses(marcher, hip):
    import pmarch
    from pmarch._lib import settings_from
    sc = scenes.lattice_cube(0, width=64, height=48)
    cam = sc.cameras[0]
    marcher.set_frame(61, sc.xyz, sc.h)
    L = pmarch.load()
    big_w, big_h = 2048, 1025                                        # 2,099,200 > 2^21 pixels
    d = _alloc(hip, 16 * big_w * big_h)
    try:
        inv = np.ascontiguousarray(cam.inv_pv, np.float32)
        eye = np.ascontiguousarray(cam.eye, np.float32)
        s = settings_from(sc.settings())
        rc = L.pm_march_device_noalloc(marcher.ctx, 61, C.byref(s), inv.ctypes.data, eye.ctypes.data, big_w, big_h,
                                       d, d, None)
        assert rc == PM_ERR_SIZE
        # the bound skip with bits that were never built for this iso
        s = settings_from(sc.settings(skip_mode=4, iso=sc.iso * 1.5))
        rc = L.pm_march_device_noalloc(marcher.ctx, 61, C.byref(s), inv.ctypes.data, eye.ctypes.data, 64, 48,
                                       d, d, None)
        assert rc == PM_ERR_STATE
        # bits built for this iso with 8x8x8 sub-boxes, asked for with 4x4x4 ones
        marcher.bound_grid(61, sc.iso, 0)
        s = settings_from(sc.settings(skip_mode=4, bound_boxes=1))
        rc = L.pm_march_device_noalloc(marcher.ctx, 61, C.byref(s), inv.ctypes.data, eye.ctypes.data, 64, 48,
                                       d, d, None)
        assert rc == PM_ERR_STATE
        assert marcher.query() == 0                                  # nothing was queued
    finally:
        hip.hipFree(d)
