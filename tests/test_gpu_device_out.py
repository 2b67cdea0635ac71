"""Outputs straight into device memory the caller owns (pm_march_async_device,
the HIP side of handing images to the renderer without host copies): the
images equal those of pm_march_async in both modes (configured with t, and
the reference march reading a device depth buffer, without t), and
pm_download refuses afterwards."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from parity import depth_in_front
from pmarch import PmError, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _alloc(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p.value


def _to_host(hip, ptr, shape, dtype=np.float32):
    out = np.zeros(shape, dtype)
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0        # device to host
    return out


def test_device_outputs_config_and_reference(marcher, hip):
    sc = scenes.lattice_cube(0, width=203, height=77)
    cam = sc.cameras[0]
    W, H = cam.width, cam.height
    print("device outputs 203x77", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    bufs = [_alloc(hip, 16 * W * H), _alloc(hip, 16 * W * H), _alloc(hip, 4 * W * H), _alloc(hip, 4 * W * H)]
    try:
        st = sc.settings(skip_mode=4)
        want = marcher.march(0, st, cam.inv_pv, cam.eye, W, H)
        stats = marcher.march_device(0, st, cam.inv_pv, cam.eye, W, H, bufs[0], bufs[1], bufs[2])
        assert np.array_equal(_to_host(hip, bufs[0], (H, W, 4)), want[0])
        assert np.array_equal(_to_host(hip, bufs[1], (H, W, 4)), want[1])
        assert np.array_equal(_to_host(hip, bufs[2], (H, W)), want[2])
        assert stats["hits"] == want[3]["hits"]
        with pytest.raises(PmError):
            marcher.download(W, H)
        # reference march from a device depth buffer, no t
        G = O.Grid(sc.xyz, sc.h)
        depth = depth_in_front(cam, G.bmin, G.bmax, 0.05)
        assert hip.hipMemcpy(bufs[3], depth.ctypes.data, depth.nbytes, 1) == 0   # host to device
        rs = dict(mode=1, max_steps=128, step=0.009, iso=sc.iso, normal_mode=1, skip_mode=2, compensated=0)
        want = marcher.march(0, rs, cam.inv_pv, cam.eye, W, H, depth=depth)
        marcher.march_device(0, rs, cam.inv_pv, cam.eye, W, H, bufs[0], bufs[1], 0, d_depth=bufs[3])
        got_p, got_n = _to_host(hip, bufs[0], (H, W, 4)), _to_host(hip, bufs[1], (H, W, 4))
        assert np.array_equal(got_p, want[0])
        assert np.array_equal(got_n, want[1], equal_nan=True)
    finally:
        hip.hipDeviceSynchronize()
        for b in bufs:
            hip.hipFree(b)
