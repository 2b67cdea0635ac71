"""Outputs straight into device memory the caller owns (pm_march_async_device,
the HIP side of handing images to the renderer without host copies): the
images equal those of pm_march_async in both modes (configured with t, and
the reference march reading a device depth buffer, without t), and
pm_download refuses afterwards.  The depth smoothing on device memory
(pm_depth_smooth_device) equals the host-image pass and the oracle.  Host
addresses, buffers shorter than the image and overlapping images are refused
before any launch (each such case first asks HIP itself that the buffer is what
the case says, so no kernel ever sees it)."""
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
    lib.hipPointerGetAttributes.argtypes = [C.c_void_p, C.c_void_p]
    lib.hipMemGetAddressRange.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.c_void_p]
    lib.hipGetLastError.argtypes = []
    return lib


class _PtrAttr(C.Structure):
    _fields_ = [("type", C.c_int), ("device", C.c_int), ("devicePointer", C.c_void_p),
                ("hostPointer", C.c_void_p), ("isManaged", C.c_int), ("allocationFlags", C.c_uint)]


def _not_device(hip, ptr):
    """HIP itself says ptr is not device memory (2: hipMemoryTypeDevice)."""
    a = _PtrAttr()
    rc = hip.hipPointerGetAttributes(C.byref(a), ptr)
    hip.hipGetLastError()
    return rc != 0 or a.type not in (2, 3)


def _extent(hip, ptr):
    base, size = C.c_void_p(), C.c_size_t()
    assert hip.hipMemGetAddressRange(C.byref(base), C.byref(size), ptr) == 0
    return base.value + size.value - ptr


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


def test_device_outputs_refused(marcher, hip):
    sc = scenes.lattice_cube(0, width=1024, height=1024)      # images far longer than a 256-byte buffer
    cam = sc.cameras[0]
    W, H = cam.width, cam.height
    marcher.set_frame(0, sc.xyz, sc.h)
    st = sc.settings(skip_mode=4)
    host = np.zeros((H, W, 4), np.float32)
    small = _alloc(hip, 256)
    big = [_alloc(hip, 16 * W * H), _alloc(hip, 16 * W * H)]
    try:
        assert _not_device(hip, host.ctypes.data)
        with pytest.raises(PmError):
            marcher.march_device(0, st, cam.inv_pv, cam.eye, W, H, host.ctypes.data, big[1])
        assert _extent(hip, small) < 16 * W * H
        with pytest.raises(PmError):
            marcher.march_device(0, st, cam.inv_pv, cam.eye, W, H, big[0], small)
        # the context still marches afterwards
        marcher.march_device(0, st, cam.inv_pv, cam.eye, W, H, big[0], big[1])
        assert np.array_equal(_to_host(hip, big[0], (H, W, 4)), marcher.march(0, st, cam.inv_pv, cam.eye, W, H)[0])
    finally:
        hip.hipDeviceSynchronize()
        for b in [small] + big:
            hip.hipFree(b)


@pytest.mark.parametrize("W,H,N", [(203, 77, 8), (97, 131, 31), (64, 48, 0)])
def test_depth_smooth_device(marcher, hip, W, H, N):
    rng = np.random.default_rng(W + N)
    depth = rng.uniform(0.2, 1.0, (H, W)).astype(np.float32)
    depth[: H // 3, : W // 4] = 1.0
    print(f"device depth smooth {W}x{H} N={N}", flush=True)
    d_in, d_out = _alloc(hip, depth.nbytes), _alloc(hip, depth.nbytes)
    try:
        assert hip.hipMemcpy(d_in, depth.ctypes.data, depth.nbytes, 1) == 0
        marcher.depth_smooth_device(d_in, d_out, W, H, N)
        hip.hipDeviceSynchronize()
        got = _to_host(hip, d_out, (H, W))
        assert np.array_equal(got, marcher.depth_smooth(depth, N))
        assert np.array_equal(got, O.depth_smooth(depth, N))
    finally:
        hip.hipDeviceSynchronize()
        hip.hipFree(d_in)
        hip.hipFree(d_out)


def test_depth_smooth_device_refused(marcher, hip):
    W, H = 1024, 1024
    host = np.zeros((H, W), np.float32)
    d_in, d_out, small = _alloc(hip, 4 * W * H), _alloc(hip, 4 * W * H), _alloc(hip, 256)
    try:
        assert _not_device(hip, host.ctypes.data)
        with pytest.raises(PmError):
            marcher.depth_smooth_device(host.ctypes.data, d_out, W, H, 8)
        assert _extent(hip, small) < 4 * W * H
        with pytest.raises(PmError):
            marcher.depth_smooth_device(d_in, small, W, H, 8)
        with pytest.raises(PmError):
            marcher.depth_smooth_device(d_in, d_in, W, H, 8)            # in place: windows overlap writes
        with pytest.raises(PmError):
            marcher.depth_smooth_device(d_in, d_out, W, H, 32)
    finally:
        hip.hipDeviceSynchronize()
        for b in (d_in, d_out, small):
            hip.hipFree(b)
