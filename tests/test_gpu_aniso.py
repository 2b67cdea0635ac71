"""GPU parity of the anisotropic march (PerPixel_Anisotropic, RayMarcher.cpp:
346-423) through the C ABI against the CPU oracle on the same seeded inputs:
WPCA's G, the density (FP32 and FP64) and the configured and reference
marches, bit for bit, with equal counters."""
import numpy as np
import pytest

import oracle as O
from parity import depth_in_front
from pmarch import scenes

pytestmark = pytest.mark.gpu


def _queries(G, xyz, rng, n):
    near = xyz[rng.integers(0, len(xyz), n // 2)] + rng.uniform(-0.15, 0.15, (n // 2, 3))
    box = rng.uniform(G.bmin - 0.3, G.bmax + 0.3, (n - n // 2, 3))
    return np.concatenate([near, box]).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1])
def test_aniso_matrix_bitexact(marcher, seed):
    sc = scenes.lattice_cube(seed)
    print(f"aniso matrix lattice_cube seed={seed}", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    st = scenes.aniso_settings(sc)
    s = O.make_settings(**st)
    q = _queries(G, sc.xyz, np.random.default_rng(seed + 20), 3000)
    Gg, dg, ng = marcher.aniso_matrix(0, st, q)
    for i in range(0, len(q), 3):
        n, Go, do = G.aniso_G(s, q[i])
        assert ng[i] == n, i
        if n:
            assert np.array_equal(Gg[i], Go) and dg[i] == np.float32(do), (i, Gg[i], Go)
    assert (ng == 0).sum() > 100 and (ng > 200).sum() > 100


@pytest.mark.parametrize("compensated", [1, 0])
def test_density_aniso_bitexact(marcher, compensated):
    sc = scenes.dam_break_aniso(1, nx=24, ny=40, nz=24)
    print(f"aniso density dam 24x40x24 compensated={compensated}", flush=True)
    marcher.set_frame(1, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    st = scenes.aniso_settings(sc, compensated=compensated)
    s = O.make_settings(**st)
    rng = np.random.default_rng(31)
    q = _queries(G, sc.xyz, rng, 6000)
    rg = marcher.density_aniso(1, st, q)
    ro = G.density_aniso(s, q)
    assert np.array_equal(rg, ro), f"max diff {np.abs(rg - ro).max()}"
    assert (ro > 1.0).sum() > 500
    q64 = q[:1500].astype(np.float64) + rng.uniform(-1e-4, 1e-4, (1500, 3))
    assert np.array_equal(marcher.density_aniso64(1, st, q64), G.density_aniso64(s, q64))


def test_density_aniso_capped_list(marcher):
    """More than 8192 extended neighbours: the capped walk on both sides."""
    rng = np.random.default_rng(5)
    v = rng.normal(size=(12000, 3))
    v = v / np.linalg.norm(v, axis=1)[:, None] * rng.uniform(0, 1, (12000, 1)) ** (1 / 3) * 0.08
    xyz = (0.5 + v).astype(np.float32)
    print("aniso density capped blob", flush=True)
    marcher.set_frame(2, xyz, 0.1)
    G = O.Grid(xyz, 0.1)
    sc = scenes.tiny_scene(xyz, 0.1, eye=(1.5, 0.5, 0.5), target=(0.5, 0.5, 0.5))
    st = scenes.aniso_settings(sc)
    s = O.make_settings(**st)
    q = (0.5 + rng.uniform(-0.2, 0.2, (400, 3))).astype(np.float32)
    _, _, ng = marcher.aniso_matrix(2, st, q)
    assert (ng > 8192).sum() > 50
    assert np.array_equal(marcher.density_aniso(2, st, q), G.density_aniso(s, q))


def _march_pair(m, frame, sc, st, cam, row_step, depth=None):
    pos, nrm, t, stats = m.march(frame, st, cam.inv_pv, cam.eye, cam.width, cam.height, depth=depth)
    G = O.Grid(sc.xyz, sc.h)
    opos, onrm, ot, ocnt = O.march(G, O.make_settings(**st), cam.inv_pv, cam.eye, cam.width, cam.height,
                                   depth=depth, row_step=row_step)
    rows = slice(0, None, row_step)
    assert np.array_equal(pos[rows], opos[rows]), "positions differ"
    assert np.array_equal(nrm[rows], onrm[rows]), "normals differ"
    assert np.array_equal(t[rows], ot[rows]), "t differs"
    if row_step == 1:
        for k in ("rays", "rays_in_aabb", "hits", "samples", "cand", "march_samples"):
            assert stats[k] == ocnt[k], (k, stats[k], ocnt[k])
    return pos, stats


@pytest.mark.parametrize("cam_i,normal_mode,skip", [(0, 0, 0), (1, 0, 1), (0, 1, 1), (0, 0, 4)])
def test_march_aniso_config_bitexact(marcher, cam_i, normal_mode, skip):
    sc = scenes.lattice_cube(0)
    cam = sc.cameras[cam_i]
    print(f"aniso config march lattice_cube cam={cam_i} normal={normal_mode} skip={skip}", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    st = scenes.aniso_settings(sc, normal_mode=normal_mode, skip_mode=skip)
    pos, stats = _march_pair(marcher, 0, sc, st, cam, 1)
    assert stats["hits"] > 50000


def test_march_aniso_schedules_agree(marcher):
    """One ray per lane to its end (schedule 0) and the persistent march +
    hit-list finish (schedule 1) give the same images and counters."""
    sc = scenes.lattice_cube(1)
    cam = sc.cameras[0]
    print("aniso config march lattice_cube schedules 0/1", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    out = []
    for v in (0, 1):
        st = scenes.aniso_settings(sc, skip_mode=1, schedule=v)
        out.append(marcher.march(0, st, cam.inv_pv, cam.eye, cam.width, cam.height))
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a, b)
    for k in ("rays", "rays_in_aabb", "hits", "samples", "cand", "march_samples"):
        assert out[0][3][k] == out[1][3][k], k


def test_march_aniso_dam_bitexact(marcher):
    """The dam-break column scaled to h = 0.1, 480x270, whole image."""
    sc = scenes.dam_break_aniso(0, width=480, height=270)
    cam = sc.cameras[0]
    print("aniso config march dam_break_aniso 480x270", flush=True)
    marcher.set_frame(3, sc.xyz, sc.h)
    for skip, schedule in ((1, 0), (4, 0), (4, 1)):
        st = scenes.aniso_settings(sc, skip_mode=skip, schedule=schedule)
        pos, stats = _march_pair(marcher, 3, sc, st, cam, 1)
        assert stats["hits"] > 10000


def test_march_aniso_odd_size_and_misses(marcher):
    """A ragged image (not a multiple of the 8x8 tile), rays that miss the box."""
    sc = scenes.lattice_cube(2, width=203, height=77)
    cam = sc.cameras[1]
    print("aniso config march 203x77", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    st = scenes.aniso_settings(sc, compensated=0, final_fp64=0, bisect_iters=5)
    _march_pair(marcher, 0, sc, st, cam, 1)


def test_march_aniso_reference_bitexact(marcher):
    """PerPixel_Anisotropic itself: depth-buffer start, MaxSteps, DensityGrid
    skip (box flag), plain sums, normal = normalize(sum gradW)."""
    sc = scenes.lattice_cube(1)
    cam = sc.cameras[0]
    print("aniso reference march lattice_cube", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    depth = depth_in_front(cam, G.bmin, G.bmax, 0.05)
    st = scenes.aniso_settings(sc, mode=1, max_steps=128, step=0.009, normal_mode=1, skip_mode=2, compensated=0)
    pos, stats = _march_pair(marcher, 0, sc, st, cam, 1, depth=depth)
    assert stats["hits"] > 50000


def test_aniso_settings_rejected(marcher):
    sc = scenes.lattice_cube(0)
    marcher.set_frame(0, sc.xyz, sc.h)
    cam = sc.cameras[0]
    for bad in (dict(learned=1), dict(k_s=0.0), dict(k_r=-1.0), dict(ext_mult=0.5), dict(ext_mult=2.5)):
        st = scenes.aniso_settings(sc, **bad)
        with pytest.raises(Exception):
            marcher.march(0, st, cam.inv_pv, cam.eye, 64, 48)


@pytest.mark.parametrize("case", ["coincident", "mixed"])
def test_aniso_degenerate_neighbourhoods(marcher, case):
    """Singular covariances: three coincident particles (C = 0: the clamp gives
    Sigma = 0 and G is NaN, so the density is never >= iso there) and a
    coincident pair next to a collinear row (rank-deficient C, clamped)."""
    if case == "coincident":
        xyz = np.array([[0.5, 0.5, 0.5]] * 3 + [[2.0, 2.0, 2.0]], np.float32)
    else:
        xyz = np.array([[0.5, 0.5, 0.5]] * 2 + [[0.5, 0.5, 0.55]] + [[0.5 + 0.02 * k, 0.5, 0.7] for k in range(5)],
                       np.float32)
    sc = scenes.tiny_scene(xyz, 0.1, eye=(0.5, 0.5, -0.5), target=(0.5, 0.5, 0.55), width=48, height=48)
    print(f"aniso degenerate {case}", flush=True)
    marcher.set_frame(6, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    st = scenes.aniso_settings(sc)
    s = O.make_settings(**st)
    rng = np.random.default_rng(9)
    q = (np.array([0.5, 0.5, 0.55]) + rng.uniform(-0.12, 0.12, (300, 3))).astype(np.float32)
    q[:3] = [[0.5, 0.5, 0.5], [0.5, 0.5, 0.52], [0.54, 0.5, 0.7]]
    Gg, dg, ng = marcher.aniso_matrix(6, st, q)
    for i in range(len(q)):
        n, Go, do = G.aniso_G(s, q[i])
        assert ng[i] == n
        if n:
            assert np.array_equal(Gg[i], Go, equal_nan=True) and np.array_equal(dg[i], np.float32(do), equal_nan=True)
    assert np.array_equal(marcher.density_aniso(6, st, q), G.density_aniso(s, q), equal_nan=True)
    _march_pair(marcher, 6, sc, st, sc.cameras[0], 1)
