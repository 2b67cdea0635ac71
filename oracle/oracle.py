"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product path (bachelor-thesis_amd/) never
imports this module.

Each wrapper names the reference function its C counterpart restates; the C
source (oracle/pm_oracle.c) carries the file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

RESTING_LATTICE_SUM = 3.1415061532636885
ZERO_ISO = float(np.float32(2.0 ** -100))
"""A sub-box bound below this means no particle within h of the sub-box."""
"""sum_j W(x_j)/sigma over a perfect lattice of spacing h/2 (27 non-zero terms)."""


class Settings(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("max_steps", C.c_int32),
        ("step", C.c_float),
        ("iso", C.c_float),
        ("bisect_iters", C.c_int32),
        ("final_fp64", C.c_int32),
        ("normal_eps", C.c_float),
        ("normal_mode", C.c_int32),
        ("skip_mode", C.c_int32),
        ("compensated", C.c_int32),
        ("anisotropic", C.c_int32),
        ("k_n", C.c_float),
        ("k_r", C.c_float),
        ("k_s", C.c_float),
        ("n_eps", C.c_int32),
        ("ext_mult", C.c_float),
        ("bound_boxes", C.c_int32),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("rays_in_aabb", C.c_uint64),
        ("hits", C.c_uint64),
        ("samples", C.c_uint64),
        ("cand", C.c_uint64),
        ("march_samples", C.c_uint64),
        ("march_cand", C.c_uint64),
        ("cand_stencil", C.c_uint64),
        ("cand27", C.c_uint64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class GridStruct(C.Structure):
    _fields_ = [
        ("h", C.c_float), ("inv_h", C.c_float),
        ("pmin", C.c_float * 3), ("pmax", C.c_float * 3),
        ("bmin", C.c_float * 3), ("bmax", C.c_float * 3),
        ("origin", C.c_int32 * 3), ("dims", C.c_int32 * 3),
        ("ncells", C.c_uint64), ("n", C.c_uint64),
        ("cell_start", C.POINTER(C.c_uint32)),
        ("sorted", C.POINTER(C.c_float)),
        ("sorted_index", C.POINTER(C.c_uint32)),
        ("dg_dims", C.c_int32 * 3),
        ("dg_flag_box", C.POINTER(C.c_uint8)),
        ("dg_flag_cons", C.POINTER(C.c_uint8)),
        ("dg_count_box", C.POINTER(C.c_uint32)),
        ("empty27", C.POINTER(C.c_uint8)),
        ("bound", C.POINTER(C.c_uint64)),
        ("bound_iso", C.c_float),
        ("bound_boxes", C.c_int32),
        ("inv_hf", C.c_float),
        ("forigin", C.c_int32 * 3),
        ("fdims", C.c_int32 * 3),
        ("fncells", C.c_uint64),
        ("fcell_start", C.POINTER(C.c_uint32)),
        ("fsorted", C.POINTER(C.c_float)),
    ]


_lib = None


def _default_threads() -> int:
    """Worker threads of the oracle's marches: the machine's cores less one,
    at most 16 (a shared machine's share of cores per GPU)."""
    return max(1, min(16, (os.cpu_count() or 2) - 1))


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        fp = np.ctypeslib.ndpointer
        L.or_W_f.restype = C.c_float
        L.or_W_f.argtypes = [C.c_float, C.c_float]
        L.or_W_d.restype = C.c_double
        L.or_W_d.argtypes = [C.c_double, C.c_double]
        L.or_sigma_f.restype = C.c_float
        L.or_sigma_f.argtypes = [C.c_float]
        L.or_gradW_f.argtypes = [fp(np.float32, flags="C"), C.c_float, fp(np.float32, flags="C")]
        L.or_key.restype = C.c_int32
        L.or_key.argtypes = [C.c_float, C.c_float]
        L.or_grid_struct_size.restype = C.c_uint64
        L.or_grid_build.argtypes = [fp(np.float32, flags="C"), C.c_uint64, C.c_float, C.POINTER(GridStruct)]
        L.or_grid_free.argtypes = [C.POINTER(GridStruct)]
        L.or_bound_build.argtypes = [C.POINTER(GridStruct), C.c_float, C.c_int32]
        L.or_neighbors_fine.restype = C.c_uint32
        L.or_neighbors_fine.argtypes = [C.POINTER(GridStruct), fp(np.float32, flags="C"), fp(np.float32, flags="C"),
                                        C.c_uint32, C.POINTER(C.c_uint32)]
        L.or_stencil_cand.restype = C.c_uint32
        L.or_stencil_cand.argtypes = [C.POINTER(GridStruct), fp(np.float32, flags="C")]
        L.or_cand27.restype = C.c_uint32
        L.or_cand27.argtypes = [C.POINTER(GridStruct), fp(np.float32, flags="C")]
        L.or_cdiff_union.restype = C.c_int
        L.or_cdiff_union.argtypes = [C.POINTER(GridStruct), fp(np.float32, flags="C"), C.c_float,
                                     C.POINTER(C.c_uint32), C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
        L.or_neighbors_brute.restype = C.c_uint32
        L.or_neighbors_brute.argtypes = [fp(np.float32, flags="C"), C.c_uint64, C.c_float,
                                         fp(np.float32, flags="C"), fp(np.uint32, flags="C"), C.c_uint32]
        L.or_neighbors_grid.restype = C.c_uint32
        L.or_neighbors_grid.argtypes = [C.POINTER(GridStruct), fp(np.float32, flags="C"),
                                        fp(np.uint32, flags="C"), C.c_uint32, C.POINTER(C.c_uint32)]
        L.or_neighbors_hash.restype = C.c_uint32
        L.or_neighbors_hash.argtypes = [fp(np.float32, flags="C"), C.c_uint64, C.c_float,
                                        fp(np.float32, flags="C"), fp(np.uint32, flags="C"), C.c_uint32]
        L.or_density_batch_f.argtypes = [C.POINTER(GridStruct), fp(np.float32, flags="C"), C.c_uint64,
                                         C.c_int, fp(np.float32, flags="C"), C.c_void_p]
        L.or_density_batch_d.argtypes = [C.POINTER(GridStruct), fp(np.float64, flags="C"), C.c_uint64,
                                         fp(np.float64, flags="C")]
        L.or_march.argtypes = [C.POINTER(GridStruct), C.POINTER(Settings), fp(np.float32, flags="C"),
                               fp(np.float32, flags="C"), C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                               C.c_int, fp(np.float32, flags="C"), fp(np.float32, flags="C"), C.c_void_p,
                               C.POINTER(Counters)]
        L.or_march_f64.argtypes = [C.POINTER(GridStruct), C.POINTER(Settings), fp(np.float32, flags="C"),
                                   fp(np.float32, flags="C"), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                   fp(np.float64, flags="C"), fp(np.float64, flags="C"), C.c_void_p]
        L.or_hash_build.restype = C.c_void_p
        L.or_hash_build.argtypes = [fp(np.float32, flags="C"), C.c_uint64, C.c_float]
        L.or_hash_free.argtypes = [C.c_void_p]
        L.or_march_hashed.argtypes = [C.c_void_p, C.POINTER(GridStruct), C.POINTER(Settings),
                                      fp(np.float32, flags="C"), fp(np.float32, flags="C"), C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.c_int, fp(np.float32, flags="C"),
                                      fp(np.float32, flags="C"), C.c_void_p, C.POINTER(Counters)]
        L.or_density_aniso_f.restype = C.c_float
        L.or_density_aniso_f.argtypes = [C.POINTER(GridStruct), C.POINTER(Settings), fp(np.float32, flags="C"),
                                         C.c_int, C.POINTER(C.c_uint32)]
        L.or_density_aniso_d.restype = C.c_double
        L.or_density_aniso_d.argtypes = [C.POINTER(GridStruct), C.POINTER(Settings), fp(np.float64, flags="C")]
        L.or_gradW_aniso_f.argtypes = [C.POINTER(GridStruct), C.POINTER(Settings), fp(np.float32, flags="C"),
                                       fp(np.float32, flags="C")]
        L.or_aniso_G_f.restype = C.c_uint32
        L.or_aniso_G_f.argtypes = [C.POINTER(GridStruct), C.POINTER(Settings), fp(np.float32, flags="C"),
                                   fp(np.float32, flags="C"), C.POINTER(C.c_float)]
        L.or_aniso_G_d.restype = C.c_uint32
        L.or_aniso_G_d.argtypes = [C.POINTER(GridStruct), C.POINTER(Settings), fp(np.float64, flags="C"),
                                   fp(np.float64, flags="C"), C.POINTER(C.c_double)]
        L.or_eig3_f.argtypes = [fp(np.float32, flags="C")] * 3
        L.or_eig3_d.argtypes = [fp(np.float64, flags="C")] * 3
        L.or_neighbors_ext.restype = C.c_uint32
        L.or_neighbors_ext.argtypes = [C.POINTER(GridStruct), fp(np.float32, flags="C"), C.c_float,
                                       fp(np.float32, flags="C"), C.c_uint32, C.POINTER(C.c_uint32)]
        L.or_gauss_kernel.argtypes = [C.c_int32, fp(np.float32, flags="C")]
        L.or_cos_half_pi.argtypes = [C.c_float]
        L.or_cos_half_pi.restype = C.c_float
        L.or_depth_sprites.argtypes = [fp(np.float32, flags="C"), C.c_uint64] + [fp(np.float32, flags="C")] * 4 + \
            [C.c_float, C.c_uint32, C.c_uint32, fp(np.float32, flags="C")]
        L.or_depth_smooth.argtypes = [fp(np.float32, flags="C"), C.c_uint32, C.c_uint32, C.c_int32,
                                      fp(np.float32, flags="C")]
        L.or_ray.argtypes = [fp(np.float32, flags="C"), fp(np.float32, flags="C"), C.c_uint32, C.c_uint32,
                             C.c_uint32, fp(np.float32, flags="C"), fp(np.float32, flags="C")]
        L.or_slab.argtypes = [fp(np.float32, flags="C")] * 5
        assert L.or_grid_struct_size() == C.sizeof(GridStruct), "GridStruct layout mismatch"
        _lib = L
    return _lib


def sigma(h: float) -> float:
    return float(lib().or_sigma_f(h))


def W(r2: float, h: float) -> float:
    """CubicSplineKernel::W (Kernel.cpp:16-32) on |r|^2, FP32."""
    return float(lib().or_W_f(r2, h))


def W64(r2: float, h: float) -> float:
    return float(lib().or_W_d(r2, h))


def gradW(r, h: float) -> np.ndarray:
    """CubicSplineKernel::gradW (Kernel.cpp:34-52), FP32, with its extra 1/|r|."""
    out = np.zeros(3, np.float32)
    lib().or_gradW_f(np.ascontiguousarray(r, np.float32), h, out)
    return out


def key(x: float, h: float) -> int:
    return int(lib().or_key(np.float32(x), np.float32(1.0) / np.float32(h)))


class Grid:
    """Dense uniform grid (cell = h) with CompactNSearch-style keys."""

    def __init__(self, xyz: np.ndarray, h: float):
        self.xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        self.h = float(np.float32(h))
        self.g = GridStruct()
        rc = lib().or_grid_build(self.xyz.reshape(-1), self.xyz.shape[0], self.h, C.byref(self.g))
        if rc != 0:
            raise RuntimeError(f"or_grid_build failed ({rc})")

    def __del__(self):
        try:
            lib().or_grid_free(C.byref(self.g))
        except Exception:
            pass

    @property
    def dims(self):
        return tuple(self.g.dims)

    @property
    def origin(self):
        return tuple(self.g.origin)

    @property
    def bmin(self):
        return np.array(self.g.bmin[:], np.float32)

    @property
    def bmax(self):
        return np.array(self.g.bmax[:], np.float32)

    def cell_start(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.g.cell_start, shape=(self.g.ncells + 1,)).copy()

    def sorted_xyz(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.g.sorted, shape=(self.g.n, 4)).copy()

    def sorted_index(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.g.sorted_index, shape=(self.g.n,)).copy()

    def densitygrid(self):
        n = int(np.prod(self.g.dg_dims[:]))
        return (tuple(self.g.dg_dims),
                np.ctypeslib.as_array(self.g.dg_count_box, shape=(n,)).copy(),
                np.ctypeslib.as_array(self.g.dg_flag_box, shape=(n,)).copy(),
                np.ctypeslib.as_array(self.g.dg_flag_cons, shape=(n,)).copy())

    def bound(self, iso: float, boxes: int = 0) -> np.ndarray:
        """Sub-box bound bits for iso and bound_boxes, (ncells, 8) words: 512 bits
        per cell (builds them when the grid holds none for iso and boxes)."""
        iso = float(np.float32(iso))
        if not self.g.bound or float(np.float32(self.g.bound_iso)) != iso or self.g.bound_boxes != boxes:
            if lib().or_bound_build(C.byref(self.g), iso, int(boxes)) != 0:
                raise RuntimeError("or_bound_build failed")
        return np.ctypeslib.as_array(self.g.bound, shape=(self.g.ncells * 8,)).reshape(-1, 8).copy()

    def empty27(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.g.empty27, shape=(self.g.ncells,)).copy()

    def neighbors(self, q) -> tuple[np.ndarray, int]:
        q = np.ascontiguousarray(q, np.float32)
        out = np.zeros(1 << 16, np.uint32)
        cand = C.c_uint32(0)
        n = lib().or_neighbors_grid(C.byref(self.g), q, out, out.size, C.byref(cand))
        return out[:n].copy(), int(cand.value)

    def march_grid(self):
        """(origin, dims, cell_start, sorted) of the march grid (cells of h/2)."""
        n = int(self.g.fncells)
        return (tuple(self.g.forigin), tuple(self.g.fdims),
                np.ctypeslib.as_array(self.g.fcell_start, shape=(n + 1,)).copy(),
                np.ctypeslib.as_array(self.g.fsorted, shape=(self.g.n, 4)).copy())

    def neighbors_fine(self, q) -> tuple[np.ndarray, int]:
        """Positions of the particles within h of q, in the march grid's order,
        and the number of candidates the trimmed fine stencil visits."""
        q = np.ascontiguousarray(q, np.float32)
        out = np.zeros(3 * (1 << 15), np.float32)
        cand = C.c_uint32(0)
        n = lib().or_neighbors_fine(C.byref(self.g), q, out, 1 << 15, C.byref(cand))
        return out[:3 * n].reshape(-1, 3).copy(), int(cand.value)

    def stencil_cand(self, q) -> np.ndarray:
        """Per point: candidates of its own trimmed fine stencil (no particle read)."""
        q = np.ascontiguousarray(np.asarray(q, np.float32).reshape(-1, 3))
        return np.array([lib().or_stencil_cand(C.byref(self.g), p) for p in q], np.uint64)

    def cand27(self, q) -> np.ndarray:
        """Per point: particles stored in the 27 search cells around its cell."""
        q = np.ascontiguousarray(np.asarray(q, np.float32).reshape(-1, 3))
        return np.array([lib().or_cand27(C.byref(self.g), p) for p in q], np.uint64)

    def cdiff_union(self, q, eps):
        """The fused central-difference walk around q: (candidate count,
        positions of the candidates within (h + eps)(1 + 2^-10) of q), or None
        when its cells leave the grid."""
        q = np.ascontiguousarray(q, np.float32)
        cap = 1 << 15
        out = np.zeros(3 * cap, np.float32)
        cand, n = C.c_uint32(0), C.c_uint32(0)
        ok = lib().or_cdiff_union(C.byref(self.g), q, float(np.float32(eps)), C.byref(cand), out.ctypes.data, cap,
                                  C.byref(n))
        if not ok:
            return None
        assert n.value <= cap
        return int(cand.value), out[:3 * n.value].reshape(-1, 3).copy()

    def density(self, q, compensated=True, with_cand=False):
        q = np.ascontiguousarray(np.asarray(q, np.float32).reshape(-1, 3))
        rho = np.zeros(q.shape[0], np.float32)
        cand = np.zeros(q.shape[0], np.uint32)
        lib().or_density_batch_f(C.byref(self.g), q.reshape(-1), q.shape[0], int(compensated), rho,
                                 cand.ctypes.data)
        return (rho, cand) if with_cand else rho

    def density_aniso(self, settings, q, compensated=None, with_cand=False):
        """Anisotropic density (RayMarcher.cpp:378-400) at each row of q, FP32."""
        q = np.ascontiguousarray(np.asarray(q, np.float32).reshape(-1, 3))
        comp = settings.compensated if compensated is None else int(compensated)
        rho = np.zeros(q.shape[0], np.float32)
        cand = np.zeros(q.shape[0], np.uint32)
        c = C.c_uint32(0)
        for i in range(q.shape[0]):
            rho[i] = lib().or_density_aniso_f(C.byref(self.g), C.byref(settings), q[i], comp, C.byref(c))
            cand[i] = c.value
        return (rho, cand) if with_cand else rho

    def density_aniso64(self, settings, q):
        q = np.ascontiguousarray(np.asarray(q, np.float64).reshape(-1, 3))
        return np.array([lib().or_density_aniso_d(C.byref(self.g), C.byref(settings), q[i])
                         for i in range(q.shape[0])], np.float64)

    def gradW_aniso(self, settings, q) -> np.ndarray:
        out = np.zeros(3, np.float32)
        lib().or_gradW_aniso_f(C.byref(self.g), C.byref(settings), np.ascontiguousarray(q, np.float32), out)
        return out

    def aniso_G(self, settings, q, fp64=False):
        """(number of extended neighbours, G (3x3), det G) of WPCA at q."""
        if fp64:
            G = np.zeros(9, np.float64)
            det = C.c_double(0.0)
            n = lib().or_aniso_G_d(C.byref(self.g), C.byref(settings), np.ascontiguousarray(q, np.float64), G,
                                   C.byref(det))
        else:
            G = np.zeros(9, np.float32)
            det = C.c_float(0.0)
            n = lib().or_aniso_G_f(C.byref(self.g), C.byref(settings), np.ascontiguousarray(q, np.float32), G,
                                   C.byref(det))
        return int(n), G.reshape(3, 3), float(det.value)

    def neighbors_ext(self, q, hx) -> tuple[np.ndarray, int]:
        """Positions of the particles with |x - q|^2 < hx^2 in the walk order, and the candidates visited."""
        q = np.ascontiguousarray(q, np.float32)
        cap = 1 << 16
        out = np.zeros(3 * cap, np.float32)
        cand = C.c_uint32(0)
        n = lib().or_neighbors_ext(C.byref(self.g), q, float(np.float32(hx)), out, cap, C.byref(cand))
        assert n <= cap
        return out[:3 * n].reshape(-1, 3).copy(), int(cand.value)

    def density64(self, q):
        q = np.ascontiguousarray(np.asarray(q, np.float64).reshape(-1, 3))
        rho = np.zeros(q.shape[0], np.float64)
        lib().or_density_batch_d(C.byref(self.g), q.reshape(-1), q.shape[0], rho)
        return rho


def eig3(c6, fp64=False):
    """Jacobi eigen-decomposition of the symmetric 3x3 matrix with lower
    triangle c6 = (c00, c10, c11, c20, c21, c22): (eigenvalues, V) with
    eigenvector k in column k of V."""
    dt = np.float64 if fp64 else np.float32
    c = np.ascontiguousarray(c6, dt)
    lam = np.zeros(3, dt)
    V = np.zeros(9, dt)
    (lib().or_eig3_d if fp64 else lib().or_eig3_f)(c, lam, V)
    return lam, V.reshape(3, 3)


def gauss_kernel(N: int) -> np.ndarray:
    """ComputeGaussKernel (GaussRenderPass.cpp:20-48): the (N+1)^2 table, [i][j]."""
    K = np.zeros((N + 1) * (N + 1), np.float32)
    assert lib().or_gauss_kernel(N, K) == 0
    return K.reshape(N + 1, N + 1)


def cos_half_pi(s) -> np.ndarray:
    """or_cos_half_pi on every element: cos(pi/2 s) as the sprite pass evaluates it."""
    s = np.ascontiguousarray(s, np.float32).reshape(-1)
    return np.array([lib().or_cos_half_pi(float(v)) for v in s], np.float32)


def depth_sprites(xyz, cam, radius: float) -> np.ndarray:
    """or_depth_sprites: the sprite depth image (H x W) of the particles xyz seen
    by cam (view, proj, inv_pv, eye as float32)."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    m = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (cam.view, cam.proj, cam.inv_pv, cam.eye)]
    out = np.zeros((cam.height, cam.width), np.float32)
    assert lib().or_depth_sprites(xyz.reshape(-1), xyz.shape[0], m[0], m[1], m[2], m[3], float(radius),
                                  cam.width, cam.height, out.reshape(-1)) == 0
    return out


def depth_smooth(depth, N: int) -> np.ndarray:
    """gauss.frag over a row-major depth image, clamp-to-edge, Spread = 1."""
    depth = np.ascontiguousarray(depth, np.float32)
    H, W = depth.shape
    out = np.zeros_like(depth)
    assert lib().or_depth_smooth(depth.reshape(-1), W, H, N, out.reshape(-1)) == 0
    return out


def neighbors_brute(xyz, h, q) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    out = np.zeros(1 << 16, np.uint32)
    n = lib().or_neighbors_brute(xyz, xyz.size // 3, h, np.ascontiguousarray(q, np.float32), out, out.size)
    return out[:n].copy()


def neighbors_hash(xyz, h, q) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    out = np.zeros(1 << 16, np.uint32)
    n = lib().or_neighbors_hash(xyz, xyz.size // 3, h, np.ascontiguousarray(q, np.float32), out, out.size)
    return out[:n].copy()


def ray(inv_pv, cam, W: int, H: int, x: int, y: int):
    """Ray of pixel (x, y): origin = camera, direction = normalize(unproject(clip, 1) - camera)."""
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().or_ray(np.ascontiguousarray(inv_pv, np.float32).reshape(16), np.ascontiguousarray(cam, np.float32),
                 W, H, y * W + x, o, d)
    return o, d


def slab(o, d, bmin, bmax):
    tnf = np.zeros(2, np.float32)
    ok = lib().or_slab(*(np.ascontiguousarray(v, np.float32) for v in (o, d, bmin, bmax)), tnf)
    return bool(ok), float(tnf[0]), float(tnf[1])


def make_settings(**kw) -> Settings:
    s = Settings()
    defaults = dict(mode=0, max_steps=0, step=0.0, iso=0.0, bisect_iters=8, final_fp64=1,
                    normal_eps=0.0, normal_mode=0, skip_mode=0, compensated=1,
                    anisotropic=0, k_n=0.5, k_r=2.0, k_s=2000.0, n_eps=1, ext_mult=2.0, bound_boxes=0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(s, k, v)
    return s


def march(grid: Grid, settings: Settings, inv_pv, cam, W: int, H: int, depth=None, row_step=1,
          threads=None, want_t=True):
    """Oracle march, FP32 (configured or reference mode, selected by settings.mode)."""
    if threads is None:
        threads = _default_threads()
    inv_pv = np.ascontiguousarray(inv_pv, np.float32).reshape(16)
    cam = np.ascontiguousarray(cam, np.float32).reshape(3)
    pos = np.zeros((H, W, 4), np.float32)
    nrm = np.zeros((H, W, 4), np.float32)
    t = np.full((H, W), -1.0, np.float32) if want_t else None
    dptr = None
    if depth is not None:
        depth = np.ascontiguousarray(depth, np.float32)
        dptr = depth.ctypes.data
    if settings.skip_mode == 4:
        # the anisotropic density skips the sub-boxes with no particle within h
        grid.bound(ZERO_ISO if settings.anisotropic else settings.iso, settings.bound_boxes)
    cnt = Counters()
    rc = lib().or_march(C.byref(grid.g), C.byref(settings), inv_pv, cam, W, H, dptr, row_step, threads,
                        pos.reshape(-1), nrm.reshape(-1), t.ctypes.data if want_t else None, C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"or_march failed ({rc})")
    return pos, nrm, t, cnt.as_dict()


def march64(grid: Grid, settings: Settings, inv_pv, cam, W: int, H: int, row_step=1, threads=None):
    """FP64 configured march: the ground truth."""
    if threads is None:
        threads = _default_threads()
    inv_pv = np.ascontiguousarray(inv_pv, np.float32).reshape(16)
    cam = np.ascontiguousarray(cam, np.float32).reshape(3)
    pos = np.zeros((H, W, 4), np.float64)
    nrm = np.zeros((H, W, 4), np.float64)
    t = np.full((H, W), -1.0, np.float64)
    rc = lib().or_march_f64(C.byref(grid.g), C.byref(settings), inv_pv, cam, W, H, row_step, threads,
                            pos.reshape(-1), nrm.reshape(-1), t.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"or_march_f64 failed ({rc})")
    return pos, nrm, t


class HashedSearch:
    """CompactNSearch-style hashed cell table (27 lookups per query)."""

    def __init__(self, xyz, h):
        self.xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
        self.h = float(np.float32(h))
        self.p = lib().or_hash_build(self.xyz, self.xyz.size // 3, self.h)

    def __del__(self):
        try:
            lib().or_hash_free(self.p)
        except Exception:
            pass

    def march(self, grid: Grid, settings: Settings, inv_pv, cam, W, H, row_step=1, threads=None):
        if threads is None:
            threads = _default_threads()
        inv_pv = np.ascontiguousarray(inv_pv, np.float32).reshape(16)
        cam = np.ascontiguousarray(cam, np.float32).reshape(3)
        pos = np.zeros((H, W, 4), np.float32)
        nrm = np.zeros((H, W, 4), np.float32)
        t = np.full((H, W), -1.0, np.float32)
        cnt = Counters()
        rc = lib().or_march_hashed(self.p, C.byref(grid.g), C.byref(settings), inv_pv, cam, W, H, row_step,
                                   threads, pos.reshape(-1), nrm.reshape(-1), t.ctypes.data, C.byref(cnt))
        if rc != 0:
            raise RuntimeError(f"or_march_hashed failed ({rc})")
        return pos, nrm, t, cnt.as_dict()
