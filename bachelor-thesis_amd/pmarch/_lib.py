"""ctypes binding of libpmarch.so (the C ABI declared in include/pmarch.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  There is no
CPU fallback: if the library is missing, or no GPU is present, these wrappers
raise.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import re

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parent.parent          # bachelor-thesis_amd/
ROOT = PKG_DIR.parent
LIB_PATH = PKG_DIR / "libpmarch.so"
HEADER = ROOT / "include" / "pmarch.h"

PM_OK, PM_PENDING = 0, 1
MODE_CONFIG, MODE_REFERENCE = 0, 1
NORMAL_CDIFF, NORMAL_GRADW = 0, 1
SKIP_NONE, SKIP_EMPTY, SKIP_DG_BOX, SKIP_DG_CONS, SKIP_BOUND = 0, 1, 2, 3, 4
MLP_INPUTS, MLP_PARAMS = 24, 6337


class PmSettings(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("max_steps", C.c_int32), ("step", C.c_float), ("iso", C.c_float),
        ("bisect_iters", C.c_int32), ("final_fp64", C.c_int32), ("normal_eps", C.c_float),
        ("normal_mode", C.c_int32), ("skip_mode", C.c_int32), ("compensated", C.c_int32),
        ("learned", C.c_int32), ("schedule", C.c_int32), ("learned_alpha", C.c_float),
        ("learned_min_skip", C.c_int32), ("learned_max_skip", C.c_int32),
        ("anisotropic", C.c_int32), ("k_n", C.c_float), ("k_r", C.c_float), ("k_s", C.c_float),
        ("n_eps", C.c_int32), ("ext_mult", C.c_float), ("instrument", C.c_int32), ("learned_gate", C.c_float),
        ("bound_boxes", C.c_int32),
    ]


class PmStats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64), ("rays_in_aabb", C.c_uint64), ("hits", C.c_uint64), ("samples", C.c_uint64),
        ("cand", C.c_uint64), ("march_samples", C.c_uint64), ("mlp_evals", C.c_uint64), ("march_cand", C.c_uint64),
        ("cert_walks", C.c_uint64), ("cert_skipped", C.c_uint64), ("cand_stencil", C.c_uint64),
        ("cand27", C.c_uint64), ("reserved", C.c_uint64), ("kernel_ms", C.c_float), ("total_ms", C.c_float),
    ]

    def as_dict(self):
        return {k: (float(getattr(self, k)) if k.endswith("_ms") else int(getattr(self, k)))
                for k, _ in self._fields_ if k != "reserved"}


class PmFrameInfo(C.Structure):
    _fields_ = [
        ("n", C.c_uint64), ("ncells", C.c_uint64), ("origin", C.c_int32 * 3), ("dims", C.c_int32 * 3),
        ("h", C.c_float), ("pmin", C.c_float * 3), ("pmax", C.c_float * 3), ("bmin", C.c_float * 3),
        ("bmax", C.c_float * 3), ("build_ms", C.c_float), ("forigin", C.c_int32 * 3), ("fdims", C.c_int32 * 3),
        ("fncells", C.c_uint64), ("bound_ms", C.c_float), ("bound_iso", C.c_float),
    ]


class PmView(C.Structure):
    _fields_ = [("frame", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32), ("inv_pv", C.c_float * 16),
                ("cam", C.c_float * 3)]


_lib = None


def header_functions() -> list[str]:
    """Names of the functions include/pmarch.h declares."""
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", txt)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"libpmarch.so not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(str(LIB_PATH))
    vp, u32, u64, i32, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_float
    L.pm_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pm_destroy.argtypes = [vp]
    L.pm_destroy.restype = None
    L.pm_last_error.argtypes = [vp]
    L.pm_last_error.restype = C.c_char_p
    L.pm_set_frame.argtypes = [vp, i32, vp, u64, f32]
    L.pm_set_frame_device.argtypes = [vp, i32, vp, u64, f32]
    L.pm_frame_info_get.argtypes = [vp, i32, C.POINTER(PmFrameInfo)]
    L.pm_get_grid.argtypes = [vp, i32, vp, vp, vp]
    L.pm_march_async.argtypes = [vp, i32, C.POINTER(PmSettings), vp, vp, u32, u32, vp, vp, vp, vp]
    L.pm_query.argtypes = [vp]
    L.pm_wait.argtypes = [vp]
    L.pm_download.argtypes = [vp, vp, vp, vp]
    L.pm_get_stats.argtypes = [vp, C.POINTER(PmStats)]
    L.pm_neighbor_counts.argtypes = [vp, i32, vp, u64, vp, vp]
    L.pm_density.argtypes = [vp, i32, vp, u64, C.c_int, vp]
    L.pm_density64.argtypes = [vp, i32, vp, u64, vp]
    L.pm_density_aniso.argtypes = [vp, i32, C.POINTER(PmSettings), vp, u64, vp]
    L.pm_density_aniso64.argtypes = [vp, i32, C.POINTER(PmSettings), vp, u64, vp]
    L.pm_aniso_matrix.argtypes = [vp, i32, C.POINTER(PmSettings), vp, u64, vp, vp, vp]
    L.pm_bgeo_read.argtypes = [C.c_char_p, vp, u64, C.POINTER(u64)]
    L.pm_gauss_kernel.argtypes = [i32, vp]
    L.pm_march_async_device.argtypes = [vp, i32, C.POINTER(PmSettings), vp, vp, u32, u32, vp, vp, vp, vp]
    L.pm_depth_smooth.argtypes = [vp, vp, u32, u32, i32, vp]
    L.pm_depth_smooth_device.argtypes = [vp, vp, u32, u32, i32, vp, vp]
    L.pm_depth_sprites.argtypes = [vp, i32, vp, vp, vp, vp, f32, u32, u32, vp]
    L.pm_depth_sprites_device.argtypes = [vp, i32, vp, vp, vp, vp, f32, u32, u32, vp]
    L.pm_load_bgeo.argtypes = [vp, i32, C.c_char_p, f32]
    L.pm_set_frame_stream.argtypes = [vp, i32, vp, u64, f32]
    L.pm_stream_copy_gbps.argtypes = [vp, u64, C.c_int, C.POINTER(C.c_float)]
    L.pm_get_skip_grids.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.pm_mlp_set.argtypes = [vp, vp]
    L.pm_mlp_get.argtypes = [vp, vp]
    L.pm_mlp_forward.argtypes = [vp, vp, u64, vp]
    L.pm_mlp_grad.argtypes = [vp, vp, vp, u64, f32, vp, vp]
    L.pm_dataset_clear.argtypes = [vp, u64]
    L.pm_trace_append.argtypes = [vp, i32, C.POINTER(PmSettings), vp, vp, u32, u32, u32, u32, u32, u64, C.POINTER(u64)]
    L.pm_dataset_get.argtypes = [vp, u64, u64, vp, vp, vp, vp]
    L.pm_mlp_train.argtypes = [vp, C.c_int, u32, f32, f32, u64, vp]
    L.pm_get_bound_grid.argtypes = [vp, i32, f32, i32, vp]
    L.pm_get_march_grid.argtypes = [vp, i32, vp, vp]
    L.pm_batch_set.argtypes = [vp, C.POINTER(PmSettings), vp, u32, C.c_int]
    L.pm_batch_run.argtypes = [vp]
    L.pm_batch_download.argtypes = [vp, u32, vp, vp, vp]
    L.pm_march_device_noalloc.argtypes = [vp, i32, C.POINTER(PmSettings), vp, vp, u32, u32, vp, vp, vp]
    L.pm_noalloc_scratch_bytes.argtypes = []
    L.pm_noalloc_scratch_bytes.restype = u64
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def bgeo_read(path) -> np.ndarray:
    """Positions (n, 3) float32 of a BGEO V5 file (plain or gzip) through
    pm_bgeo_read -- host code of the library, no GPU needed."""
    L = load()
    n = C.c_uint64(0)
    bpath = str(path).encode()
    rc = L.pm_bgeo_read(bpath, None, 0, C.byref(n))
    if rc != PM_OK:
        raise PmError(f"pm_bgeo_read({path}) failed ({rc})")
    xyz = np.zeros((n.value, 3), np.float32)
    rc = L.pm_bgeo_read(bpath, _ptr(xyz), n.value, C.byref(n))
    if rc != PM_OK:
        raise PmError(f"pm_bgeo_read({path}) failed ({rc})")
    return xyz


def gauss_kernel(N: int) -> np.ndarray:
    """pm_gauss_kernel: ComputeGaussKernel's (N+1)^2 table, [i][j] (host code)."""
    K = np.zeros((N + 1) * (N + 1), np.float32)
    if load().pm_gauss_kernel(int(N), _ptr(K)) != PM_OK:
        raise PmError(f"pm_gauss_kernel({N}) failed")
    return K.reshape(N + 1, N + 1)


def settings_from(d: dict) -> PmSettings:
    s = PmSettings()
    for k, v in d.items():
        setattr(s, k, v)
    return s


class PmError(RuntimeError):
    pass


class Marcher:
    """Thin owner of a pm_ctx (one GPU)."""

    def __init__(self, device: int = 0):
        self.L = load()
        p = C.c_void_p()
        rc = self.L.pm_create(device, C.byref(p))
        if rc != PM_OK:
            raise PmError(f"pm_create failed ({rc}): no usable GPU")
        self.ctx = p
        self._keep = {}

    def close(self):
        if getattr(self, "ctx", None):
            self.L.pm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            raise PmError(f"{what} failed ({rc}): {self.L.pm_last_error(self.ctx).decode()}")
        return rc

    def set_frame(self, frame: int, xyz: np.ndarray, h: float):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        self._check(self.L.pm_set_frame(self.ctx, frame, _ptr(xyz), xyz.shape[0], float(h)), "pm_set_frame")
        return self.frame_info(frame)

    def set_frame_stream(self, frame: int, xyz: np.ndarray, h: float):
        """pm_set_frame_stream: builds the frame beside a march of another frame in flight."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        self._check(self.L.pm_set_frame_stream(self.ctx, frame, _ptr(xyz), xyz.shape[0], float(h)),
                    "pm_set_frame_stream")
        return self.frame_info(frame)

    def load_bgeo(self, frame: int, path, h: float):
        self._check(self.L.pm_load_bgeo(self.ctx, frame, str(path).encode(), float(h)), "pm_load_bgeo")
        return self.frame_info(frame)

    def depth_smooth(self, depth, N: int = 8):
        """GaussRenderPass's smoothing of a row-major depth image on the GPU."""
        depth = np.ascontiguousarray(depth, np.float32)
        H, W = depth.shape
        out = np.zeros_like(depth)
        self._check(self.L.pm_depth_smooth(self.ctx, _ptr(depth), W, H, int(N), _ptr(out)), "pm_depth_smooth")
        return out

    def depth_sprites(self, frame: int, cam, radius: float):
        """pm_depth_sprites: the sprite depth image (H x W, 1.0 where nothing is
        drawn) of `frame` seen by `cam` (a pmarch.camera.Camera)."""
        m = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (cam.view, cam.proj, cam.inv_pv, cam.eye)]
        out = np.zeros((cam.height, cam.width), np.float32)
        self._check(self.L.pm_depth_sprites(self.ctx, frame, _ptr(m[0]), _ptr(m[1]), _ptr(m[2]), _ptr(m[3]),
                                            float(radius), cam.width, cam.height, _ptr(out)), "pm_depth_sprites")
        return out

    def depth_sprites_device(self, frame: int, cam, radius: float, d_depth: int):
        """pm_depth_sprites_device: the sprite depths into device memory
        (integer device address), enqueued on the context stream."""
        m = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (cam.view, cam.proj, cam.inv_pv, cam.eye)]
        self._keep_sprites = m
        self._check(self.L.pm_depth_sprites_device(self.ctx, frame, _ptr(m[0]), _ptr(m[1]), _ptr(m[2]), _ptr(m[3]),
                                                   float(radius), cam.width, cam.height, d_depth),
                    "pm_depth_sprites_device")

    def depth_smooth_device(self, d_depth: int, d_out: int, W: int, H: int, N: int = 8, stream: int = 0):
        """pm_depth_smooth_device: enqueue the smoothing of device memory
        (integer device addresses) on `stream` (an integer hipStream_t handle,
        0 for the null stream); returns without waiting."""
        self._check(self.L.pm_depth_smooth_device(self.ctx, d_depth, W, H, int(N), d_out, stream or None),
                    "pm_depth_smooth_device")

    def frame_info(self, frame: int) -> PmFrameInfo:
        fi = PmFrameInfo()
        self._check(self.L.pm_frame_info_get(self.ctx, frame, C.byref(fi)), "pm_frame_info_get")
        return fi

    def grid(self, frame: int):
        fi = self.frame_info(frame)
        cs = np.zeros(fi.ncells + 1, np.uint32)
        srt = np.zeros((fi.n, 4), np.float32)
        idx = np.zeros(fi.n, np.uint32)
        self._check(self.L.pm_get_grid(self.ctx, frame, _ptr(cs), _ptr(srt), _ptr(idx)), "pm_get_grid")
        return fi, cs, srt, idx

    def march_async(self, frame: int, settings: dict, inv_pv, cam, W: int, H: int, depth=None,
                    outputs: bool = True, want_t: bool = True):
        s = settings_from(settings)
        inv_pv = np.ascontiguousarray(inv_pv, np.float32).reshape(16)
        cam = np.ascontiguousarray(cam, np.float32).reshape(3)
        pos = nrm = t = None
        if outputs:
            pos = np.zeros((H, W, 4), np.float32)
            nrm = np.zeros((H, W, 4), np.float32)
            t = np.zeros((H, W), np.float32) if want_t else None
        if depth is not None:
            depth = np.ascontiguousarray(depth, np.float32)
        self._keep = dict(s=s, inv_pv=inv_pv, cam=cam, depth=depth)
        self._check(self.L.pm_march_async(self.ctx, frame, C.byref(s), _ptr(inv_pv), _ptr(cam), W, H, _ptr(depth),
                                          _ptr(pos), _ptr(nrm), _ptr(t)), "pm_march_async")
        return pos, nrm, t

    def march_device(self, frame: int, settings: dict, inv_pv, cam, W: int, H: int, d_pos: int, d_nrm: int,
                     d_t: int = 0, d_depth: int = 0):
        """pm_march_async_device + wait: the images go into the caller's device
        memory (integer device addresses)."""
        s = settings_from(settings)
        inv_pv = np.ascontiguousarray(inv_pv, np.float32).reshape(16)
        cam = np.ascontiguousarray(cam, np.float32).reshape(3)
        self._keep = dict(s=s, inv_pv=inv_pv, cam=cam)
        self._check(self.L.pm_march_async_device(self.ctx, frame, C.byref(s), _ptr(inv_pv), _ptr(cam), W, H,
                                                 d_depth or None, d_pos, d_nrm, d_t or None), "pm_march_async_device")
        self.wait()
        return self.stats()

    def march_device_noalloc(self, frame: int, settings: dict, inv_pv, cam, W: int, H: int, d_pos: int,
                             d_nrm: int, d_t: int = 0):
        """pm_march_device_noalloc + wait: the march into the caller's device
        memory with the library's static scratch (no allocation)."""
        s = settings_from(settings)
        inv_pv = np.ascontiguousarray(inv_pv, np.float32).reshape(16)
        cam = np.ascontiguousarray(cam, np.float32).reshape(3)
        self._keep = dict(s=s, inv_pv=inv_pv, cam=cam)
        self._check(self.L.pm_march_device_noalloc(self.ctx, frame, C.byref(s), _ptr(inv_pv), _ptr(cam), W, H,
                                                   d_pos, d_nrm, d_t or None), "pm_march_device_noalloc")
        self.wait()
        return self.stats()

    def query(self) -> int:
        return self._check(self.L.pm_query(self.ctx), "pm_query")

    def wait(self):
        self._check(self.L.pm_wait(self.ctx), "pm_wait")

    def march(self, frame: int, settings: dict, inv_pv, cam, W: int, H: int, depth=None, want_t=True):
        pos, nrm, t = self.march_async(frame, settings, inv_pv, cam, W, H, depth, True, want_t)
        self.wait()
        return pos, nrm, t, self.stats()

    def stats(self) -> dict:
        st = PmStats()
        self._check(self.L.pm_get_stats(self.ctx, C.byref(st)), "pm_get_stats")
        return st.as_dict()

    def march_grid(self, frame: int):
        fi = self.frame_info(frame)
        cs = np.zeros(fi.fncells + 1, np.uint32)
        srt = np.zeros((fi.n, 4), np.float32)
        self._check(self.L.pm_get_march_grid(self.ctx, frame, _ptr(cs), _ptr(srt)), "pm_get_march_grid")
        return fi, cs, srt

    def bound_grid(self, frame: int, iso: float, boxes: int = 0) -> np.ndarray:
        """The frame's sub-box bound bits for iso and bound_boxes: (ncells, 8) words
        of 512 bits per cell."""
        fi = self.frame_info(frame)
        bits = np.zeros((fi.ncells, 8), np.uint64)
        self._check(self.L.pm_get_bound_grid(self.ctx, frame, float(iso), int(boxes), _ptr(bits)), "pm_get_bound_grid")
        return bits

    def batch_set(self, settings: dict, views, use_graph: bool = True):
        """views: iterable of (frame, W, H, inv_pv, cam)."""
        vs = list(views)
        arr = (PmView * len(vs))()
        for i, (frame, W, H, inv_pv, cam) in enumerate(vs):
            arr[i].frame, arr[i].width, arr[i].height = int(frame), int(W), int(H)
            arr[i].inv_pv[:] = [float(x) for x in np.asarray(inv_pv, np.float32).reshape(16)]
            arr[i].cam[:] = [float(x) for x in np.asarray(cam, np.float32).reshape(3)]
        s = settings_from(settings)
        self._check(self.L.pm_batch_set(self.ctx, C.byref(s), C.cast(arr, C.c_void_p), len(vs), int(bool(use_graph))),
                    "pm_batch_set")
        self._batch_sizes = [(int(v[1]), int(v[2])) for v in vs]

    def batch_run(self):
        self._check(self.L.pm_batch_run(self.ctx), "pm_batch_run")

    def batch_download(self, view: int):
        W, H = self._batch_sizes[view]
        pos = np.zeros((H, W, 4), np.float32)
        nrm = np.zeros((H, W, 4), np.float32)
        t = np.zeros((H, W), np.float32)
        self._check(self.L.pm_batch_download(self.ctx, view, _ptr(pos), _ptr(nrm), _ptr(t)), "pm_batch_download")
        return pos, nrm, t

    def download(self, W, H):
        pos = np.zeros((H, W, 4), np.float32)
        nrm = np.zeros((H, W, 4), np.float32)
        t = np.zeros((H, W), np.float32)
        self._check(self.L.pm_download(self.ctx, _ptr(pos), _ptr(nrm), _ptr(t)), "pm_download")
        return pos, nrm, t

    def skip_grids(self, frame: int):
        fi = self.frame_info(frame)
        dims = np.zeros(3, np.int32)
        self._check(self.L.pm_get_skip_grids(self.ctx, frame, None, _ptr(dims), None, None, None), "pm_get_skip_grids")
        n = int(np.prod(dims))
        e27 = np.zeros(fi.ncells, np.uint8)
        fb = np.zeros(n, np.uint8)
        fc = np.zeros(n, np.uint8)
        cb = np.zeros(n, np.uint32)
        self._check(self.L.pm_get_skip_grids(self.ctx, frame, _ptr(e27), _ptr(dims), _ptr(fb), _ptr(fc), _ptr(cb)),
                    "pm_get_skip_grids")
        return e27, tuple(int(v) for v in dims), cb, fb, fc

    # ---- learned step proposal ----
    def mlp_set(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        assert p.size == MLP_PARAMS
        self._check(self.L.pm_mlp_set(self.ctx, _ptr(p)), "pm_mlp_set")

    def mlp_get(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        self._check(self.L.pm_mlp_get(self.ctx, _ptr(p)), "pm_mlp_get")
        return p

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        y = np.zeros(x.shape[0], np.float32)
        self._check(self.L.pm_mlp_forward(self.ctx, _ptr(x), x.shape[0], _ptr(y)), "pm_mlp_forward")
        return y

    def mlp_grad(self, x, t, w_over=4.0):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.L.pm_mlp_grad(self.ctx, _ptr(x), _ptr(t), x.shape[0], float(w_over), _ptr(loss), _ptr(g)),
                    "pm_mlp_grad")
        return float(loss[0]), g

    def dataset_clear(self, capacity: int):
        self._check(self.L.pm_dataset_clear(self.ctx, int(capacity)), "pm_dataset_clear")

    def trace_append(self, frame, settings, inv_pv, cam, W, H, ray_stride=16, k_stride=3, seed=0, max_new_rows=0) -> int:
        s = settings_from(settings)
        inv_pv = np.ascontiguousarray(inv_pv, np.float32).reshape(16)
        cam = np.ascontiguousarray(cam, np.float32).reshape(3)
        tot = C.c_uint64(0)
        self._check(self.L.pm_trace_append(self.ctx, frame, C.byref(s), _ptr(inv_pv), _ptr(cam), W, H, ray_stride,
                                           k_stride, seed, int(max_new_rows), C.byref(tot)), "pm_trace_append")
        return int(tot.value)

    def dataset_get(self, offset, n):
        f = np.zeros((n, MLP_INPUTS), np.float32)
        y = np.zeros(n, np.float32)
        hit = np.zeros(n, np.uint8)
        meta = np.zeros((n, 2), np.uint32)
        self._check(self.L.pm_dataset_get(self.ctx, offset, n, _ptr(f), _ptr(y), _ptr(hit), _ptr(meta)),
                    "pm_dataset_get")
        return f, y, hit, meta

    def mlp_train(self, epochs, batch=4096, lr=2e-3, w_over=4.0, seed=0):
        losses = np.zeros(epochs, np.float32)
        self._check(self.L.pm_mlp_train(self.ctx, epochs, batch, float(lr), float(w_over), seed, _ptr(losses)),
                    "pm_mlp_train")
        return losses

    def stream_copy_gbps(self, nbytes: int, iters: int = 20) -> float:
        g = C.c_float(0.0)
        self._check(self.L.pm_stream_copy_gbps(self.ctx, nbytes, iters, C.byref(g)), "pm_stream_copy_gbps")
        return float(g.value)

    def neighbor_counts(self, frame: int, q):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        cnt = np.zeros(q.shape[0], np.uint32)
        cand = np.zeros(q.shape[0], np.uint32)
        self._check(self.L.pm_neighbor_counts(self.ctx, frame, _ptr(q), q.shape[0], _ptr(cnt), _ptr(cand)),
                    "pm_neighbor_counts")
        return cnt, cand

    def density(self, frame: int, q, compensated=True):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        rho = np.zeros(q.shape[0], np.float32)
        self._check(self.L.pm_density(self.ctx, frame, _ptr(q), q.shape[0], int(compensated), _ptr(rho)), "pm_density")
        return rho

    def density64(self, frame: int, q):
        q = np.ascontiguousarray(q, np.float64).reshape(-1, 3)
        rho = np.zeros(q.shape[0], np.float64)
        self._check(self.L.pm_density64(self.ctx, frame, _ptr(q), q.shape[0], _ptr(rho)), "pm_density64")
        return rho

    def density_aniso(self, frame: int, settings: dict, q):
        """Anisotropic density (RayMarcher.cpp:378-400) at query points, FP32."""
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        rho = np.zeros(q.shape[0], np.float32)
        s = settings_from(settings)
        self._check(self.L.pm_density_aniso(self.ctx, frame, C.byref(s), _ptr(q), q.shape[0], _ptr(rho)),
                    "pm_density_aniso")
        return rho

    def density_aniso64(self, frame: int, settings: dict, q):
        q = np.ascontiguousarray(q, np.float64).reshape(-1, 3)
        rho = np.zeros(q.shape[0], np.float64)
        s = settings_from(settings)
        self._check(self.L.pm_density_aniso64(self.ctx, frame, C.byref(s), _ptr(q), q.shape[0], _ptr(rho)),
                    "pm_density_aniso64")
        return rho

    def aniso_matrix(self, frame: int, settings: dict, q):
        """WPCA at query points: (G [n,3,3], det G [n], extended-neighbour counts [n])."""
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        n = q.shape[0]
        G = np.zeros((n, 9), np.float32)
        det = np.zeros(n, np.float32)
        cnt = np.zeros(n, np.uint32)
        s = settings_from(settings)
        self._check(self.L.pm_aniso_matrix(self.ctx, frame, C.byref(s), _ptr(q), n, _ptr(G), _ptr(det), _ptr(cnt)),
                    "pm_aniso_matrix")
        return G.reshape(n, 3, 3), det, cnt
