"""Pins of the oracle's march (RayMarcher.cpp:256-344 and the configured march):
closed-form ray geometry, single/two-particle roots, normals, FP32 vs FP64,
dense vs hashed search, reference-mode rules -- no GPU."""
import math

import numpy as np
import pytest

import oracle as O
from parity import assert_within, compare, depth_in_front, extent
from pmarch import camera as pc
from pmarch import scenes


def test_pixel_rays_closed_form():
    """Pixel-corner convention (x = i*2/W - 1, y = j*2/H - 1) through glm LH / y-flipped
    projection: the ray of clip (cx, cy) is f + cx*tan*aspect*s - cy*tan*u in camera axes."""
    eye, target = np.array([0.3, 1.2, -2.0]), np.array([0.5, 0.4, 0.2])
    W, H = 64, 48
    cam = pc.make_camera(eye, target, W, H, fovy_deg=60.0)
    f = (target - eye) / np.linalg.norm(target - eye)
    s = np.cross([0, 1, 0], f)
    s /= np.linalg.norm(s)
    u = np.cross(f, s)
    t = math.tan(math.radians(30.0))
    for (x, y) in [(W // 2, H // 2), (0, 0), (W - 1, 0), (17, 33), (W - 1, H - 1)]:
        o, d = O.ray(cam.inv_pv, cam.eye, W, H, x, y)
        cx, cy = x * 2.0 / W - 1.0, y * 2.0 / H - 1.0
        e = f + cx * t * (W / H) * s - cy * t * u
        e /= np.linalg.norm(e)
        assert np.allclose(o, eye.astype(np.float32))
        assert angle_deg(d, e) < 2e-4, (x, y, d, e)
    _, dc = O.ray(cam.inv_pv, cam.eye, W, H, W // 2, H // 2)
    assert np.dot(dc, f) > 1 - 1e-7          # the centre pixel looks down the view axis


def test_slab_closed_form():
    ok, tn, tf = O.slab([0, 0, -5], [0, 0, 1], [-1, -1, -1], [1, 1, 1])
    assert ok and tn == 4.0 and tf == 6.0
    ok, tn, tf = O.slab([0, 0, 0], [0, 0, 1], [-1, -1, -1], [1, 1, 1])    # inside: start at 0
    assert ok and tn == 0.0 and tf == 1.0
    ok, _, _ = O.slab([2, 0, -5], [0, 0, 1], [-1, -1, -1], [1, 1, 1])     # parallel, outside
    assert not ok
    ok, _, _ = O.slab([0, 0, 5], [0, 0, 1], [-1, -1, -1], [1, 1, 1])      # box behind
    assert not ok


def angle_deg(a, b):
    """Angle between two directions, well conditioned near zero (atan2 of |a x b|, a.b)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.degrees(np.arctan2(np.linalg.norm(np.cross(a, b)), np.dot(a, b))))


def _axis_scene(points, h, dist=2.0, W=64, H=48):
    """Camera on -z looking at the origin; with even W and H the pixel (W/2, H/2) has clip
    coordinates (0, 0) under the pixel-corner convention, so its ray is the z axis."""
    return scenes.tiny_scene(points, h, eye=(0.0, 0.0, -dist), target=(0.0, 0.0, 0.0), width=W, height=H)


@pytest.mark.parametrize("iso_frac", [0.05, 0.2, 0.6, 0.9])
def test_single_particle_root(iso_frac):
    """One particle at the origin: rho(z) = W(|z|), the crossing is at q* with
    2(1-q)^3 = iso/sigma (iso <= sigma/4) or 6(q^3-q^2)+1 = iso/sigma."""
    h = 0.125
    sc = _axis_scene([[0.0, 0.0, 0.0]], h)
    sig = O.sigma(h)
    iso = float(np.float32(iso_frac * sig))
    st = O.make_settings(**sc.settings(iso=iso))
    G = O.Grid(sc.xyz, h)
    c = sc.cameras[0]
    x, y = c.width // 2, c.height // 2
    a = iso / sig
    if a <= 0.25:
        q = 1.0 - (a / 2.0) ** (1.0 / 3.0)
    else:
        roots = np.roots([6.0, -6.0, 0.0, 1.0 - a])
        q = float([r.real for r in roots if abs(r.imag) < 1e-12 and 0 <= r.real <= 0.5][0])
    t_expect = 2.0 - q * h
    res = sc.step / 256.0
    pos, nrm, t, _ = O.march(G, st, c.inv_pv, c.eye, c.width, c.height)
    assert abs(t[y, x] - t_expect) <= res / 2 + 1e-6, (t[y, x], t_expect)
    p64, n64, t64 = O.march64(G, st, c.inv_pv, c.eye, c.width, c.height)
    assert abs(t64[y, x] - t_expect) <= res / 2 + 1e-12
    # radial normal, pointing back to the camera
    assert np.allclose(nrm[y, x, :3], [0, 0, -1], atol=1e-4)
    hit = pos[..., 3] > 0
    rad = pos[hit][:, :3] / np.linalg.norm(pos[hit][:, :3], axis=1, keepdims=True)
    assert np.all((rad * nrm[hit][:, :3]).sum(1) > 1 - 1e-4)


def test_two_particle_root_brute_force():
    h = 0.125
    pts = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, -0.06]], np.float32)
    sc = _axis_scene(pts, h)
    sig8 = 8 / (math.pi * h ** 3)
    iso = float(np.float32(0.5 * O.sigma(h)))

    def rho(z):
        return sum(O.W64(float((z - p[2]) ** 2), h) for p in pts.astype(np.float64))

    zs = np.linspace(-0.06 - h, 0.0, 200001)
    r = np.array([rho(z) for z in zs])
    k = int(np.argmax(r >= iso))
    lo, hi = zs[k - 1], zs[k]
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if rho(mid) < iso else (lo, mid)
    t_expect = 2.0 + lo
    G = O.Grid(sc.xyz, h)
    st = O.make_settings(**sc.settings(iso=iso))
    c = sc.cameras[0]
    _, _, t64 = O.march64(G, st, c.inv_pv, c.eye, c.width, c.height)
    _, _, t32, _ = O.march(G, st, c.inv_pv, c.eye, c.width, c.height)
    y, x = c.height // 2, c.width // 2
    assert abs(t64[y, x] - t_expect) <= sc.step / 512 + 1e-12
    assert abs(t32[y, x] - t_expect) <= sc.step / 256 + 1e-6
    assert sig8 > 0


def test_planar_face_normal_by_symmetry():
    """Perfect lattice slab; the centre ray hits the z face head-on above a lattice
    column: by symmetry the normal is the face normal."""
    h = 0.125
    s = h / 2
    i = np.arange(-8, 9)
    k = np.arange(0, 6)
    g = np.stack(np.meshgrid(i, i, k, indexing="ij"), -1).reshape(-1, 3) * s
    sc = _axis_scene(g.astype(np.float32), h, W=32, H=24)
    G = O.Grid(sc.xyz, h)
    st = O.make_settings(**sc.settings())
    c = sc.cameras[0]
    pos, nrm, t, _ = O.march(G, st, c.inv_pv, c.eye, c.width, c.height)
    y, x = c.height // 2, c.width // 2
    assert pos[y, x, 3] == 1.0
    assert angle_deg(nrm[y, x, :3], [0, 0, -1]) < 1e-3
    _, n64, _ = O.march64(G, st, c.inv_pv, c.eye, c.width, c.height)
    assert np.allclose(n64[y, x, :3], [0, 0, -1], atol=1e-12)


def test_fp32_vs_fp64_lattice_cube():
    sc = scenes.lattice_cube(0, width=160, height=120)
    c = sc.cameras[0]
    G = O.Grid(sc.xyz, sc.h)
    st = O.make_settings(**sc.settings())
    pos, nrm, t, cnt = O.march(G, st, c.inv_pv, c.eye, c.width, c.height)
    p64, n64, t64 = O.march64(G, st, c.inv_pv, c.eye, c.width, c.height)
    m = compare(pos, nrm, t, p64, n64, t64)
    assert_within(m, extent(sc.xyz))
    assert cnt["hits"] > 5000
    # every miss is exactly (0,0,0,0) in both images
    miss = pos[..., 3] == 0
    assert np.all(pos[miss] == 0) and np.all(nrm[miss] == 0) and np.all(t[miss] == -1)


def test_hashed_search_matches_grid_march():
    """The reference-style march (27 h-cells from a hashed table, summed in
    h-cell order) against the grid march (fine-cell order over the trimmed
    5x5x5 fine stencil): the same samples and hits, images equal to within the
    last bits of the density sums, far fewer candidates for the fine stencil."""
    sc = scenes.lattice_cube(1, width=96, height=72)
    c = sc.cameras[1]
    G = O.Grid(sc.xyz, sc.h)
    st = O.make_settings(**sc.settings())
    a = O.march(G, st, c.inv_pv, c.eye, c.width, c.height)
    b = O.HashedSearch(sc.xyz, sc.h).march(G, st, c.inv_pv, c.eye, c.width, c.height)
    m = compare(*a[:3], *b[:3])
    assert_within(m, extent(sc.xyz))
    same = np.mean(np.all(a[0] == b[0], axis=-1))
    print("identical pixels", same, "candidates fine", a[3]["cand"], "hashed h-cells", b[3]["cand"])
    assert same > 0.99
    for k in ("rays", "rays_in_aabb", "hits", "samples", "march_samples"):
        assert a[3][k] == b[3][k], k
    assert a[3]["cand"] < 0.5 * b[3]["cand"]


def test_thread_count_does_not_change_results():
    sc = scenes.lattice_cube(2, width=80, height=60)
    c = sc.cameras[0]
    G = O.Grid(sc.xyz, sc.h)
    st = O.make_settings(**sc.settings())
    a = O.march(G, st, c.inv_pv, c.eye, c.width, c.height, threads=1)
    b = O.march(G, st, c.inv_pv, c.eye, c.width, c.height, threads=7)
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u, v)


def test_reference_mode_first_sample_rule():
    """PerPixel_Isotropic: the hit is the first sample (start + k*step) with rho >= iso;
    the previous sample is below iso; normals follow normalize(sum gradW)."""
    sc = scenes.lattice_cube(0, width=48, height=36)
    c = sc.cameras[0]
    G = O.Grid(sc.xyz, sc.h)
    depth = depth_in_front(c, G.bmin, G.bmax, 0.02)
    st = O.make_settings(mode=1, max_steps=128, step=0.009, iso=sc.iso, normal_mode=1, compensated=0)
    pos, nrm, t, cnt = O.march(G, st, c.inv_pv, c.eye, c.width, c.height, depth=depth)
    hit = pos[..., 3] > 0
    assert hit.sum() > 200
    hp = pos[hit][:, :3]
    rho = G.density(hp, compensated=False)
    assert np.all(rho >= sc.iso)
    # one step back along the ray is below iso
    d = hp - c.eye[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    prev = (hp - 0.009 * d).astype(np.float32)
    assert np.mean(G.density(prev, compensated=False) < sc.iso) > 0.99
    # the normal points outward: away from the local particle mass
    assert np.all(np.abs(np.linalg.norm(nrm[hit][:, :3], axis=1) - 1) < 1e-5)
    # misses where the depth buffer is clear
    assert np.all(pos[depth == 1.0] == 0)


def test_empty_stencil_skip_keeps_images():
    """The configured march with the empty-stencil skip gives the same image with fewer samples."""
    sc = scenes.clustered_pool(0, n_total=150_000, width=128, height=72)
    c = sc.cameras[0]
    G = O.Grid(sc.xyz, sc.h)
    a = O.march(G, O.make_settings(**sc.settings()), c.inv_pv, c.eye, c.width, c.height)
    b = O.march(G, O.make_settings(**sc.settings(skip_mode=1)), c.inv_pv, c.eye, c.width, c.height)
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u, v)
    assert b[3]["samples"] < 0.7 * a[3]["samples"] and a[3]["hits"] > 100
    # cells flagged empty really have no particle within h of any of their points
    e = G.empty27()
    dims = np.array(G.dims)
    rng = np.random.default_rng(0)
    pick = rng.choice(np.flatnonzero(e), 300, replace=False)
    z, rem = np.divmod(pick, dims[0] * dims[1])
    y, x = np.divmod(rem, dims[0])
    key = np.stack([x, y, z], -1) + np.array(G.origin)
    pts = (key + rng.uniform(0.01, 0.99, key.shape)) * sc.h
    assert np.all(G.density64(pts) == 0.0)


def test_reference_mode_densitygrid_skip():
    """Reference march with the conservative DensityGrid flag: a skip moves the ray to
    the node exit plus one step (RayMarcher.cpp:302), which shifts the sample lattice,
    so hits agree with the unskipped march to within a step for most rays; a few
    grazing rays meet a different droplet."""
    sc = scenes.clustered_pool(1, n_total=150_000, width=160, height=90)
    c = sc.cameras[0]
    G = O.Grid(sc.xyz, sc.h)
    depth = depth_in_front(c, G.bmin, G.bmax, 0.02)
    kw = dict(mode=1, max_steps=400, step=0.004, iso=sc.iso, normal_mode=1, compensated=0)
    a = O.march(G, O.make_settings(**kw), c.inv_pv, c.eye, c.width, c.height, depth=depth)
    b = O.march(G, O.make_settings(skip_mode=3, **kw), c.inv_pv, c.eye, c.width, c.height, depth=depth)
    ha, hb = a[0][..., 3] > 0, b[0][..., 3] > 0
    assert np.mean(ha == hb) > 0.99 and ha.sum() > 500
    both = ha & hb
    assert np.median(np.abs(a[2][both] - b[2][both])) <= 0.004 * 1.01
    assert b[3]["samples"] < a[3]["samples"]
