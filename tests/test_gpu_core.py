"""GPU parity: grid build, neighbour counts, density and the configured march,
through the C ABI, against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

import oracle as O
from parity import assert_within, compare, extent
from pmarch import scenes

pytestmark = pytest.mark.gpu


def _grid_equal(m, frame, xyz, h):
    fi, cs, srt, idx = m.grid(frame)
    G = O.Grid(xyz, h)
    assert tuple(fi.dims) == G.dims and tuple(fi.origin) == G.origin
    assert np.array_equal(np.array(fi.bmin[:], np.float32), G.bmin)
    assert np.array_equal(np.array(fi.bmax[:], np.float32), G.bmax)
    assert np.array_equal(cs, G.cell_start()), "cell_start differs"
    assert np.array_equal(idx, G.sorted_index()), "stored particle order differs"
    assert np.array_equal(srt, G.sorted_xyz()), "stored coordinates differ"
    # the march grid (cells of h/2)
    fi, fcs, fsrt = m.march_grid(frame)
    forg, fdims, ocs, osrt = G.march_grid()
    assert tuple(fi.forigin) == forg and tuple(fi.fdims) == fdims
    assert np.array_equal(fcs, ocs), "march-grid cell_start differs"
    assert np.array_equal(fsrt, osrt), "march-grid stored coordinates differ"
    return G


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grid_lattice_cube(marcher, seed):
    sc = scenes.lattice_cube(seed)
    print(f"grid lattice_cube seed={seed} n={sc.n}", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    _grid_equal(marcher, 0, sc.xyz, sc.h)


def test_grid_dam_break(marcher):
    sc = scenes.dam_break(0)
    print(f"grid dam_break n={sc.n}", flush=True)
    marcher.set_frame(1, sc.xyz, sc.h)
    _grid_equal(marcher, 1, sc.xyz, sc.h)


def test_grid_negative_coordinates(marcher):
    rng = np.random.default_rng(5)
    xyz = rng.uniform(-0.35, 0.2, (5000, 3)).astype(np.float32)
    xyz[:7] = [[-0.1, 0, 0], [0, -0.1, 0], [0, 0, -0.1], [-0.2, -0.2, -0.2], [0.1, 0.1, 0.1], [0, 0, 0], [-0.0, 0.05, -0.3]]
    marcher.set_frame(2, xyz, 0.1)
    _grid_equal(marcher, 2, xyz, 0.1)


def test_neighbor_counts(marcher):
    sc = scenes.lattice_cube(1)
    marcher.set_frame(0, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    rng = np.random.default_rng(3)
    q = rng.uniform(G.bmin - 0.05, G.bmax + 0.05, (4000, 3)).astype(np.float32)
    cnt, cand = marcher.neighbor_counts(0, q)
    for i in range(0, 4000, 13):
        nb, nc = G.neighbors(q[i])
        assert cnt[i] == len(nb) and cand[i] == nc, i
    # a brute-force pin on a subset
    for i in range(0, 4000, 397):
        assert cnt[i] == len(O.neighbors_brute(sc.xyz, sc.h, q[i]))


@pytest.mark.parametrize("compensated", [1, 0])
def test_density_bitexact(marcher, compensated):
    sc = scenes.dam_break(2, nx=24, ny=40, nz=24)
    marcher.set_frame(3, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    rng = np.random.default_rng(11)
    q = rng.uniform(G.bmin, G.bmax, (20000, 3)).astype(np.float32)
    rg = marcher.density(3, q, compensated=bool(compensated))
    ro = G.density(q, compensated=bool(compensated))
    assert np.array_equal(rg, ro), f"max diff {np.abs(rg - ro).max()}"
    q64 = q.astype(np.float64) + rng.uniform(-1e-4, 1e-4, q.shape)
    assert np.array_equal(marcher.density64(3, q64), G.density64(q64))


@pytest.mark.parametrize("seed,cam,schedule", [(0, 0, 0), (1, 1, 0), (2, 0, 0), (0, 0, 1), (1, 1, 1)])
def test_march_lattice_cube(marcher, seed, cam, schedule):
    sc = scenes.lattice_cube(seed)
    c = sc.cameras[cam]
    print(f"march lattice_cube seed={seed} cam={cam} schedule={schedule} {c.width}x{c.height}", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    st = sc.settings(schedule=schedule)
    pos, nrm, t, stats = marcher.march(0, st, c.inv_pv, c.eye, c.width, c.height)
    G = O.Grid(sc.xyz, sc.h)
    opos, onrm, ot, ocnt = O.march(G, O.make_settings(**st), c.inv_pv, c.eye, c.width, c.height)
    m = compare(pos, nrm, t, opos, onrm, ot)
    print(m, flush=True)
    assert_within(m, extent(sc.xyz))
    assert m["hits_b"] > 50000
    # same canonical order and no contraction on either side: identical images
    assert np.array_equal(pos, opos) and np.array_equal(nrm, onrm) and np.array_equal(t, ot)
    for k in ("rays", "rays_in_aabb", "hits", "samples", "march_cand", "march_samples"):
        assert stats[k] == ocnt[k], (k, stats[k], ocnt[k])


@pytest.mark.parametrize("kw", [dict(), dict(skip_mode=4), dict(compensated=0), dict(skip_mode=1),
                                dict(normal_mode=1), dict(final_fp64=0), dict(bisect_iters=0)])
def test_march_schedules_identical(marcher, kw):
    """The persistent march (compacted walks, lane groups, per-wave finishing)
    and the one-ray-per-lane schedule (a direct walk for every sample) give the
    same images and the same sample counts, with and without the skips, with
    plain (uncompensated) sums, gradW normals and without the FP64 step or any
    bisection; an instrumented launch gives the same images too."""
    sc = scenes.dam_break(1, nx=40, ny=80, nz=40, width=320, height=240)
    c = sc.cameras[0]
    print(f"schedules {kw}", flush=True)
    marcher.set_frame(4, sc.xyz, sc.h)
    ref = marcher.march(4, sc.settings(schedule=1, **kw), c.inv_pv, c.eye, c.width, c.height)
    got = marcher.march(4, sc.settings(**kw), c.inv_pv, c.eye, c.width, c.height)
    ins = marcher.march(4, sc.settings(instrument=1, **kw), c.inv_pv, c.eye, c.width, c.height)
    print(ref[3], got[3], sep="\n", flush=True)
    assert ref[3]["hits"] > 2000
    for a, b, i in zip(ref[:3], got[:3], ins[:3]):
        assert np.array_equal(a, b) and np.array_equal(a, i), kw
    for k in ("rays", "hits", "samples", "march_samples", "march_cand"):
        assert ref[3][k] == got[3][k] == ins[3][k], (k, kw)


def test_dense_support_overflow(marcher):
    """Samples with far more particles in their support than the 32 compaction
    slots per lane (buffer drained mid-stencil) give the image of the direct
    sum and of the oracle."""
    rng = np.random.default_rng(9)
    xyz = (rng.normal(0.0, 0.03, (60000, 3)) + 0.5).astype(np.float32)
    h = 0.02
    sc = scenes.tiny_scene(xyz, h, eye=(0.5, 0.56, 0.22), target=(0.5, 0.5, 0.5), width=160, height=120)
    c = sc.cameras[0]
    marcher.set_frame(5, sc.xyz, sc.h)
    # iso 8x the lattice value puts the surface where ~130 particles lie within h
    st = sc.settings(iso=sc.iso * 8)
    n, _ = marcher.neighbor_counts(5, np.array([[0.5, 0.5 + 0.072, 0.5]], np.float32))
    print("particles within h at the surface", n[0], flush=True)
    assert n[0] > 64
    ref = marcher.march(5, dict(st, schedule=1), c.inv_pv, c.eye, c.width, c.height)
    assert ref[3]["hits"] > 1000
    got = marcher.march(5, st, c.inv_pv, c.eye, c.width, c.height)
    for a, b in zip(ref[:3], got[:3]):
        assert np.array_equal(a, b)
    G = O.Grid(sc.xyz, sc.h)
    op, on, ot, _ = O.march(G, O.make_settings(**st), c.inv_pv, c.eye, c.width, c.height)
    assert np.array_equal(ref[0], op) and np.array_equal(ref[1], on) and np.array_equal(ref[2], ot)


def test_march_vs_fp64(marcher):
    sc = scenes.lattice_cube(0)
    c = sc.cameras[0]
    marcher.set_frame(0, sc.xyz, sc.h)
    st = sc.settings()
    pos, nrm, t, _ = marcher.march(0, st, c.inv_pv, c.eye, c.width, c.height)
    G = O.Grid(sc.xyz, sc.h)
    p64, n64, t64 = O.march64(G, O.make_settings(**st), c.inv_pv, c.eye, c.width, c.height)
    m = compare(pos, nrm, t, p64, n64, t64)
    print("vs FP64:", m, flush=True)
    assert_within(m, extent(sc.xyz))


@pytest.mark.parametrize("kw", [dict(), dict(skip_mode=4), dict(final_fp64=0, compensated=0)])
def test_instrumented_counts_match_oracle(marcher, kw):
    """An instrumented launch counts, for every density evaluation however the
    kernels took it (march walk, finish list, FP64 step, the six central
    differences), the candidates of its own fine stencil and of its 27 search
    cells: both sums equal the oracle's, sample by sample restated."""
    sc = scenes.lattice_cube(1)
    c = sc.cameras[0]
    print(f"instrumented counts lattice_cube {kw}", flush=True)
    marcher.set_frame(0, sc.xyz, sc.h)
    st = sc.settings(instrument=1, **kw)
    pos, nrm, t, stats = marcher.march(0, st, c.inv_pv, c.eye, c.width, c.height)
    G = O.Grid(sc.xyz, sc.h)
    osettings = O.make_settings(**{k: v for k, v in st.items() if k in dict(O.Settings._fields_)})
    opos, onrm, ot, ocnt = O.march(G, osettings, c.inv_pv, c.eye, c.width, c.height)
    assert np.array_equal(pos, opos) and np.array_equal(nrm, onrm)
    print(stats, ocnt, sep="\n", flush=True)
    for k in ("samples", "march_samples", "march_cand", "cand_stencil", "cand27"):
        assert stats[k] == ocnt[k], (k, stats[k], ocnt[k])
    # a sample's 27 search cells hold about 2.5x its trimmed fine stencil's candidates
    assert 1.5 * stats["cand_stencil"] < stats["cand27"]
