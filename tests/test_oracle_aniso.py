"""Pins of the oracle's anisotropic density (PerPixel_Anisotropic,
RayMarcher.cpp:346-423; WPCA :114-254; AnisotropicKernel, Kernel.cpp:55-107):
the Jacobi eigen-decomposition against LAPACK and against the reference's own
solver (Eigen's computeDirect, built from the reference's vendored headers into
oracle/_ref when the reference tree is present), the extended neighbour walk
against brute force, closed forms for symmetric and degenerate neighbourhoods
and for an isolated particle, invariances, and the FP32 march against the FP64
march.  No GPU."""
import ctypes as C
import math
import pathlib

import numpy as np
import pytest

import oracle as O
from parity import assert_within, compare, extent
from pmarch import scenes

REF_EIGEN = pathlib.Path(O.__file__).resolve().parent / "_ref" / "libeigen_ref.so"
H = np.float32(0.1)


def _settings(**kw):
    d = dict(step=0.0125, iso=1.0, normal_eps=0.003125, **scenes.ANISO_SETTINGS)
    d.update(kw)
    return O.make_settings(**d)


def _lower(A):
    return np.array([A[0, 0], A[1, 0], A[1, 1], A[2, 0], A[2, 1], A[2, 2]])


# ---------------------------------------------------------------- eigen step --

def _check_eig(A, lam, V, tol):
    A = np.asarray(A, np.float64)
    lam = np.asarray(lam, np.float64)
    V = np.asarray(V, np.float64)
    scale = max(np.abs(A).max(), 1e-30)
    assert np.abs(V.T @ V - np.eye(3)).max() < tol * 10
    assert np.abs(A @ V - V * lam).max() <= tol * scale * 10
    ref = np.linalg.eigvalsh(A)
    assert np.abs(np.sort(lam) - ref).max() <= tol * scale * 10


@pytest.mark.parametrize("fp64", [False, True])
def test_jacobi_random_symmetric(fp64):
    rng = np.random.default_rng(3)
    tol = 1e-13 if fp64 else 2e-6
    for i in range(400):
        B = rng.normal(size=(3, 3))
        A = B @ B.T * 10.0 ** rng.uniform(-8, 0) if i % 2 else (B + B.T)
        if not fp64:
            A = A.astype(np.float32).astype(np.float64)
        lam, V = O.eig3(_lower(A), fp64=fp64)
        _check_eig(A, lam, V, tol)


def test_jacobi_degenerate_cases():
    cases = [np.zeros((3, 3)), np.diag([1.0, 2.0, 3.0]), np.eye(3) * 4e-3, np.ones((3, 3)),
             np.outer([1.0, 2.0, -1.0], [1.0, 2.0, -1.0]), np.diag([2.0, 2.0, 0.0]),
             np.array([[2.0, 1e-30, 0.0], [1e-30, 1.0, 0.0], [0.0, 0.0, 1.0]])]
    for A in cases:
        lam, V = O.eig3(_lower(A).astype(np.float32))
        _check_eig(A.astype(np.float32), lam, V, 2e-6)
    lam, V = O.eig3(_lower(np.diag([1.0, 2.0, 3.0])).astype(np.float32))
    assert np.array_equal(V, np.eye(3, dtype=np.float32)) and np.array_equal(lam, [1, 2, 3])   # no rotation


def _cov_samples(G, s, qs):
    L = O.lib()
    fp = np.ctypeslib.ndpointer
    L.or_aniso_C_f.restype = C.c_uint32
    L.or_aniso_C_f.argtypes = [C.POINTER(O.GridStruct), C.POINTER(O.Settings), fp(np.float32), fp(np.float32)]
    out = []
    for q in qs:
        c = np.zeros(6, np.float32)
        n = L.or_aniso_C_f(C.byref(G.g), C.byref(s), np.ascontiguousarray(q, np.float32), c)
        if n:
            out.append((c, n))
    return out


def _G_from_C(G, s, c, n):
    L = O.lib()
    fp = np.ctypeslib.ndpointer
    L.or_aniso_G_from_C_f.argtypes = [C.POINTER(O.GridStruct), C.POINTER(O.Settings), fp(np.float32), C.c_uint32,
                                      fp(np.float32), C.POINTER(C.c_float)]
    g = np.zeros(9, np.float32)
    d = C.c_float(0.0)
    L.or_aniso_G_from_C_f(C.byref(G.g), C.byref(s), c, n, g, C.byref(d))
    return g, d.value


def _two_pass_cov(nb, q, hx):
    """RayMarcher.cpp:138-169 in FP64 with numpy: weights 1 - (|x - q| / hx)^3,
    the weighted mean m, then C = sum w (x - m)(x - m)^T / sum w."""
    nb = nb.astype(np.float64)
    d = np.linalg.norm(nb - q.astype(np.float64), axis=1)
    w = 1.0 - (d / hx) ** 3
    m = (w[:, None] * nb).sum(0) / w.sum()
    x = nb - m
    return (w[:, None, None] * x[:, :, None] * x[:, None, :]).sum(0) / w.sum()


@pytest.mark.parametrize("scene", ["lattice", "sheet", "line"])
def test_one_pass_covariance_equals_two_pass(scene):
    """WPCA's covariance about the weighted mean, taken in one walk as
    sum w u u^T / W - mu mu^T with u = x - (first neighbour), against the
    reference's two-pass form restated in FP64 over the brute-force extended
    neighbours: the FP64 oracle to 1e-12 of the largest entry, the FP32 one to
    2e-5 -- at points in the bulk, on the surface and outside, where the mean
    is far from q and a few neighbours lie close together."""
    rng = np.random.default_rng(11)
    if scene == "lattice":
        xyz = scenes.lattice_cube(3).xyz
    elif scene == "sheet":
        i = np.arange(-12, 13) * 0.05
        xyz = (np.stack(np.meshgrid(i, i, [0.0], indexing="ij"), -1).reshape(-1, 3) + 0.5)
        xyz = (xyz + rng.normal(0, 0.002, xyz.shape)).astype(np.float32)
    else:
        t = np.arange(-40, 41) * 0.01
        xyz = (0.5 + np.stack([t, 0.3 * t, -0.2 * t], -1) + rng.normal(0, 0.001, (t.size, 3))).astype(np.float32)
    G = O.Grid(xyz, H)
    s = _settings()
    hx = 2.0 * H
    L = O.lib()
    fp = np.ctypeslib.ndpointer
    L.or_aniso_C_d.restype = C.c_uint32
    L.or_aniso_C_d.argtypes = [C.POINTER(O.GridStruct), C.POINTER(O.Settings), fp(np.float64), fp(np.float64)]
    lo, hi = xyz.min(0) - 0.15, xyz.max(0) + 0.15
    qs = np.concatenate([xyz[rng.integers(0, len(xyz), 60)] + rng.normal(0, 0.03, (60, 3)),
                         rng.uniform(lo, hi, (60, 3))]).astype(np.float32)
    checked = 0
    for q in qs:
        nb = xyz[_brute_ext(xyz, q, np.float32(hx))]
        c64 = np.zeros(6)
        n = L.or_aniso_C_d(C.byref(G.g), C.byref(s), q.astype(np.float64), c64)
        assert n == len(nb)
        if n < 2:
            continue
        ref = _two_pass_cov(nb, q, hx)
        want = ref[[0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]]
        scale = np.abs(ref).max()
        assert np.abs(c64 - want).max() <= 1e-12 * scale, (scene, q)
        c32, n32 = _cov_samples(G, s, [q])[0]
        assert n32 == n
        assert np.abs(c32.astype(np.float64) - want).max() <= 2e-5 * scale, (scene, q)
        checked += 1
    assert checked > 40


@pytest.mark.skipif(not REF_EIGEN.exists(), reason="oracle/_ref not built (the reference tree is absent)")
def test_G_matches_reference_eigen_solver():
    """G and det G from the Jacobi step against Eigen::SelfAdjointEigenSolver::
    computeDirect (the reference's solver) on covariances of real samples:
    surface, bulk, edges and corners of a jittered lattice, a plane and a line."""
    R = C.CDLL(str(REF_EIGEN))
    fp = np.ctypeslib.ndpointer
    R.ref_aniso_G_from_C.argtypes = [fp(np.float32), C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                                     fp(np.float32), C.POINTER(C.c_float)]
    s = _settings()
    rng = np.random.default_rng(1)
    scenes_ = [scenes.lattice_cube(0).xyz]
    i = np.arange(-8, 9) * 0.05
    plane = np.stack(np.meshgrid(i, i, [0.0], indexing="ij"), -1).reshape(-1, 3) + 0.5
    line = np.stack([i, np.zeros_like(i), np.zeros_like(i)], -1) + 0.5
    scenes_ += [plane.astype(np.float32), line.astype(np.float32)]
    rel, drel = [], []
    for xyz in scenes_:
        G = O.Grid(xyz, H)
        lo, hi = xyz.min(0) - 0.15, xyz.max(0) + 0.15
        qs = rng.uniform(lo, hi, (3000, 3)).astype(np.float32)
        for c, n in _cov_samples(G, s, qs):
            g1, d1 = _G_from_C(G, s, c, n)
            g2 = np.zeros(9, np.float32)
            d2 = C.c_float(0.0)
            assert R.ref_aniso_G_from_C(c, n, np.float32(1) / H, 0.5, 2.0, 2000.0, 1, g2, C.byref(d2)) == 0
            rel.append(np.abs(g1 - g2).max() / np.abs(g2).max())
            drel.append(abs(d1 - d2.value) / abs(d2.value))
    rel, drel = np.array(rel), np.array(drel)
    assert len(rel) > 3000
    print(f"G rel diff: median {np.median(rel):.2e} p99 {np.quantile(rel, 0.99):.2e} max {rel.max():.2e}; "
          f"det: median {np.median(drel):.2e} max {drel.max():.2e}")
    assert np.median(rel) < 1e-6 and np.quantile(rel, 0.99) < 1e-5 and rel.max() < 1e-3
    assert np.median(drel) < 1e-6 and drel.max() < 1e-3


# --------------------------------------------------------- neighbour walk --

def _brute_ext(xyz, q, hx):
    r = xyz.astype(np.float32) - np.asarray(q, np.float32)
    r2 = (r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2]
    hx = np.float32(hx)
    return r2 < hx * hx


def test_ext_walk_matches_brute_force_in_grid_order():
    sc = scenes.lattice_cube(2)
    G = O.Grid(sc.xyz, sc.h)
    hx = np.float32(2.0) * np.float32(sc.h)
    rng = np.random.default_rng(4)
    qs = np.concatenate([rng.uniform(G.bmin - 0.25, G.bmax + 0.25, (300, 3)),
                         sc.xyz[rng.integers(0, sc.n, 100)]]).astype(np.float32)
    srt = G.sorted_xyz()[:, :3]
    rank = {tuple(p): i for i, p in enumerate(map(tuple, srt))}
    outside = 0
    for q in qs:
        got, cand = G.neighbors_ext(q, hx)
        want = sc.xyz[_brute_ext(sc.xyz, q, hx)]
        assert len(got) == len(want)
        assert cand >= len(got)
        if len(got):
            # the walk visits cells (z, y, x) in order and a cell's particles by input index:
            # its list is the stored order restricted to the neighbours
            order = [rank[tuple(p)] for p in map(tuple, got)]
            assert order == sorted(order)
        else:
            outside += 1
    assert outside > 10    # query points off the grid: the clamped walk finds nothing, no fault


# ------------------------------------------------------------ closed forms --

def test_isolated_particle_closed_form():
    """N = 1 <= N_eps: Sigma = k_n, so G = I / (h k_n) exactly and
    rho(r) = (8/pi) det G f(|G r| / h) with the cubic spline's f."""
    xyz = np.array([[0.5, 0.5, 0.5], [5.0, 5.0, 5.0]], np.float32)
    G = O.Grid(xyz, H)
    s = _settings()
    q = np.array([0.5012, 0.4995, 0.5003], np.float32)
    n, Gm, det = G.aniso_G(s, q)
    g = np.float32(np.float32(1.0) / H) * np.float32(np.float32(1.0) / np.float32(0.5))
    assert n == 1
    assert np.array_equal(Gm, np.diag([g, g, g]).astype(np.float32))
    assert det == pytest.approx(float(g) ** 3, rel=1e-6)
    for d in np.linspace(0.0, 0.006, 25):
        q = np.array([0.5 + d, 0.5, 0.5], np.float32)
        r = float(np.float32(q[0] - np.float32(0.5)))
        qq = float(g) * abs(r) / float(H)
        f = 0.0 if qq >= 1 else (6 * (qq ** 3 - qq ** 2) + 1 if qq < 0.5 else 2 * (1 - qq) ** 3)
        want = 8.0 / math.pi * float(g) ** 3 * f
        assert G.density_aniso(s, q)[0] == pytest.approx(want, rel=2e-5, abs=1e-3)


def test_isolated_particle_hit_depth_closed_form():
    """The configured anisotropic march towards an isolated particle hits the
    sphere |x - p| = R*, R* = q* h^2 k_n with (8/pi) det G f(q*) = iso, solved
    in FP64 on the closed form."""
    p = np.array([0.5, 0.5, 0.5], np.float32)
    xyz = np.array([p, [3.0, 3.0, 3.0]], np.float32)
    G = O.Grid(xyz, H)
    g = 1.0 / (0.1 * 0.5)
    A = 8.0 / math.pi * g ** 3
    iso = 1.0
    lo, hi = 0.5, 1.0          # f falls from 1/4 to 0 on [0.5, 1): iso / A lies there
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if A * 2 * (1 - mid) ** 3 >= iso: lo = mid
        else: hi = mid
    Rstar = lo * 0.1 / g
    sc = scenes.tiny_scene(xyz, 0.1, eye=(0.5, 0.5, -0.5), target=(0.5, 0.5, 0.5), width=16, height=16)
    cam = sc.cameras[0]
    s = _settings(step=Rstar / 8, iso=iso, normal_mode=0, normal_eps=Rstar / 64)
    pos, nrm, t, cnt = O.march(G, s, cam.inv_pv, cam.eye, cam.width, cam.height)
    hits = pos[..., 3] > 0
    assert hits.sum() >= 1
    x = pos[hits][:, :3].astype(np.float64)
    rad = np.linalg.norm(x - p, axis=1)
    assert np.abs(rad - Rstar).max() < 1e-4 * Rstar + 1e-6
    # the normals point away from the particle
    n = nrm[hits][:, :3].astype(np.float64)
    assert ((n * (x - p) / rad[:, None]).sum(-1) > 0.999).all()


def test_symmetric_lattice_gives_isotropic_G():
    """At a site of a perfect cubic lattice the weighted covariance is c I by
    symmetry, so G = I / (h k_s c) with c = sum w x^2 / sum w (FP64 closed sum)."""
    sp = 0.05
    i = np.arange(-10, 11) * sp
    xyz = (np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3) + 0.5).astype(np.float32)
    G = O.Grid(xyz, H)
    s = _settings()
    q = np.array([0.5, 0.5, 0.5], np.float32)
    n, Gm, det = G.aniso_G(s, q)
    r = xyz.astype(np.float64) - 0.5
    d = np.linalg.norm(r, axis=1)
    m = d < 0.2
    w = 1.0 - (d[m] / 0.2) ** 3
    c = (w * r[m, 0] ** 2).sum() / w.sum()
    assert n == m.sum()
    want = 1.0 / (0.1 * 2000.0 * c)
    assert np.abs(Gm - np.diag([want] * 3)).max() < 2e-5 * want
    assert det == pytest.approx(want ** 3, rel=6e-5)


def test_planar_sheet_clamps_normal_axis():
    """On a flat square sheet the covariance is diag(c, c, 0): the clamp lifts
    the normal eigenvalue to c / k_r, so G_zz = k_r G_xx (and, with k_r = 4,
    4 G_xx)."""
    sp = 0.05
    i = np.arange(-10, 11) * sp
    xyz = (np.stack(np.meshgrid(i, i, [0.0], indexing="ij"), -1).reshape(-1, 3) + 0.5).astype(np.float32)
    G = O.Grid(xyz, H)
    for kr in (2.0, 4.0):
        s = _settings(k_r=kr)
        n, Gm, det = G.aniso_G(s, np.array([0.5, 0.5, 0.5], np.float32))
        assert abs(Gm[0, 0] - Gm[1, 1]) < 1e-5 * Gm[0, 0]
        assert Gm[2, 2] == pytest.approx(kr * Gm[0, 0], rel=1e-5)
        assert np.abs(Gm - np.diag(np.diag(Gm))).max() < 1e-5 * Gm[0, 0]


def test_axis_permutation_invariance():
    """Cycling the axes of the scene and the point cycles G and leaves rho
    (to rounding: the walk order changes)."""
    sc = scenes.lattice_cube(1)
    s = _settings()
    G1 = O.Grid(sc.xyz, sc.h)
    G2 = O.Grid(np.ascontiguousarray(sc.xyz[:, [1, 2, 0]]), sc.h)
    rng = np.random.default_rng(8)
    qs = sc.xyz[rng.integers(0, sc.n, 200)] + rng.uniform(-0.1, 0.1, (200, 3)).astype(np.float32)
    qs = qs.astype(np.float32)
    r1 = G1.density_aniso(s, qs)
    r2 = G2.density_aniso(s, np.ascontiguousarray(qs[:, [1, 2, 0]]))
    assert np.allclose(r1, r2, rtol=2e-4, atol=1e-4)
    for q in qs[:40]:
        _, A, _ = G1.aniso_G(s, q)
        _, B, _ = G2.aniso_G(s, q[[1, 2, 0]])
        P = np.eye(3)[[1, 2, 0]]
        assert np.allclose(P @ A @ P.T, B, rtol=1e-4, atol=1e-5 * np.abs(A).max())


def test_fp32_fp64_density_agree():
    sc = scenes.lattice_cube(0)
    G = O.Grid(sc.xyz, sc.h)
    s = _settings()
    rng = np.random.default_rng(2)
    qs = (sc.xyz[rng.integers(0, sc.n, 500)] + rng.uniform(-0.12, 0.12, (500, 3))).astype(np.float32)
    r32 = G.density_aniso(s, qs).astype(np.float64)
    r64 = G.density_aniso64(s, qs.astype(np.float64))
    assert np.abs(r32 - r64).max() <= 1e-4 * np.abs(r64).max()


def test_capped_neighbour_list():
    """More than 8192 extended neighbours (RayMarcher.cpp:14, 381): WPCA and the
    density use only the first 8192 of the walk; the result is that of the
    same formula over exactly that prefix (checked in FP64), and differs from
    the formula over the whole list."""
    rng = np.random.default_rng(5)
    v = rng.normal(size=(12000, 3))
    v = v / np.linalg.norm(v, axis=1)[:, None] * rng.uniform(0, 1, (12000, 1)) ** (1 / 3) * 0.08
    xyz = (0.5 + v).astype(np.float32)
    G = O.Grid(xyz, H)
    s = _settings()
    q = np.array([0.52, 0.49, 0.5], np.float32)
    lst, _ = G.neighbors_ext(q, np.float32(2) * H)
    assert len(lst) > 8192

    def formula(nb):
        nb = nb.astype(np.float64)
        r = nb - q.astype(np.float64)
        d = np.linalg.norm(r, axis=1)
        w = 1 - (d / 0.2) ** 3
        mean = (w[:, None] * nb).sum(0) / w.sum()
        x = nb - mean
        Cm = (w[:, None, None] * x[:, :, None] * x[:, None, :]).sum(0) / w.sum()
        lam, R = np.linalg.eigh(Cm)
        lam = np.maximum(lam, lam.max() / 2.0) * 2000.0
        Gm = R @ np.diag(1 / lam) @ R.T / 0.1
        near = r[d < 0.1]
        qq = np.linalg.norm(near @ Gm.T, axis=1) / 0.1
        f = np.where(qq < 0.5, 6 * (qq ** 3 - qq ** 2) + 1, np.where(qq < 1, 2 * (1 - qq) ** 3, 0.0))
        return 8 / math.pi * np.linalg.det(Gm) * f.sum()

    rho = float(G.density_aniso(s, q)[0])
    assert rho == pytest.approx(formula(lst[:8192]), rel=1e-3)
    assert abs(rho - formula(lst)) > 1e-2 * abs(rho)


# ------------------------------------------------------------------ march --

@pytest.mark.parametrize("cam_i", [0, 1])
def test_march_fp32_within_tolerance_of_fp64(cam_i):
    """The FP32 anisotropic march (compensated sums, FP64 last bisection step)
    against the FP64 march on the 32^3 lattice at 640x480, every 4th row."""
    sc = scenes.lattice_cube(0)
    G = O.Grid(sc.xyz, sc.h)
    cam = sc.cameras[cam_i]
    s = O.make_settings(**scenes.aniso_settings(sc))
    pos, nrm, t, cnt = O.march(G, s, cam.inv_pv, cam.eye, cam.width, cam.height, row_step=4)
    p64, n64, t64 = O.march64(G, s, cam.inv_pv, cam.eye, cam.width, cam.height, row_step=4)
    rows = slice(0, None, 4)
    m = compare(pos[rows], nrm[rows], t[rows], p64[rows], n64[rows], t64[rows])
    print(m)
    assert m["hits_a"] > 5000
    assert_within(m, extent(sc.xyz))


def test_march_reference_mode_runs_with_densitygrid_skip():
    """PerPixel_Anisotropic: depth-buffer start, MaxSteps, DensityGrid skip,
    normal = normalize(sum gradW); the skip leaves the image unchanged here
    (its nodes hold no particle within h of any sample it passes)."""
    from parity import depth_in_front
    sc = scenes.lattice_cube(1)
    G = O.Grid(sc.xyz, sc.h)
    cam = sc.cameras[0]
    depth = depth_in_front(cam, G.bmin, G.bmax, 0.05)
    out = []
    for skip in (0, 3):
        s = O.make_settings(**scenes.aniso_settings(sc, mode=1, max_steps=128, step=0.009, normal_mode=1,
                                                     skip_mode=skip, compensated=0))
        out.append(O.march(G, s, cam.inv_pv, cam.eye, cam.width, cam.height, depth=depth, row_step=8))
    (p0, n0, t0, c0), (p3, n3, t3, c3) = out
    assert c0["hits"] > 1000
    assert np.array_equal(p0, p3) and np.array_equal(n0, n3)
    assert c3["march_samples"] <= c0["march_samples"]
    hit = p0[..., 3] > 0
    assert np.allclose(np.linalg.norm(n0[hit][:, :3], axis=1), 1.0, atol=1e-5)
