"""GPU parity of the skip structures (empty 27-neighbourhoods, reference DensityGrid),
of the configured march with the empty-space skip, and of the reference march
(PerPixel_Isotropic, RayMarcher.cpp:256-344) with and without its DensityGrid skip."""
import numpy as np
import pytest

import oracle as O
from parity import compare, depth_in_front
from pmarch import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    return scenes.clustered_pool(2, n_total=300_000, width=320, height=180)


def test_skip_grids_equal_oracle(marcher, pool):
    marcher.set_frame(10, pool.xyz, pool.h)
    e27, dims, cb, fb, fc = marcher.skip_grids(10)
    G = O.Grid(pool.xyz, pool.h)
    odims, ocb, ofb, ofc = G.densitygrid()
    assert np.array_equal(e27, G.empty27())
    assert dims == odims
    assert np.array_equal(cb, ocb) and np.array_equal(fb, ofb) and np.array_equal(fc, ofc)


def test_empty_skip_bitexact_and_identical(marcher, pool):
    marcher.set_frame(10, pool.xyz, pool.h)
    c = pool.cameras[0]
    G = O.Grid(pool.xyz, pool.h)
    st = pool.settings(skip_mode=1)
    pos, nrm, t, stats = marcher.march(10, st, c.inv_pv, c.eye, c.width, c.height)
    opos, onrm, ot, ocnt = O.march(G, O.make_settings(**st), c.inv_pv, c.eye, c.width, c.height)
    assert np.array_equal(pos, opos) and np.array_equal(nrm, onrm) and np.array_equal(t, ot)
    for k in ("rays", "rays_in_aabb", "hits", "samples", "cand", "march_samples"):
        assert stats[k] == ocnt[k], k
    p0, n0, t0, s0 = marcher.march(10, pool.settings(), c.inv_pv, c.eye, c.width, c.height)
    assert np.array_equal(pos, p0) and np.array_equal(nrm, n0) and np.array_equal(t, t0)
    assert stats["samples"] < s0["samples"]
    print("samples with / without skip:", stats["samples"], s0["samples"], flush=True)


@pytest.mark.parametrize("skip,normal,comp", [(0, 1, 0), (2, 1, 0), (3, 1, 0), (0, 0, 1)])
def test_reference_mode_bitexact(marcher, skip, normal, comp):
    sc = scenes.lattice_cube(1, width=320, height=240)
    c = sc.cameras[0]
    marcher.set_frame(11, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    depth = depth_in_front(c, G.bmin, G.bmax, 0.05)
    st = dict(mode=1, max_steps=128, step=0.009, iso=sc.iso, bisect_iters=0, final_fp64=0,
              normal_eps=sc.normal_eps, normal_mode=normal, skip_mode=skip, compensated=comp)
    pos, nrm, t, stats = marcher.march(11, st, c.inv_pv, c.eye, c.width, c.height, depth=depth)
    opos, onrm, ot, ocnt = O.march(G, O.make_settings(**st), c.inv_pv, c.eye, c.width, c.height, depth=depth)
    m = compare(pos, nrm, t, opos, onrm, ot)
    print(m, flush=True)
    assert m["hits_b"] > 10000
    assert np.array_equal(pos, opos) and np.array_equal(t, ot)
    # normals: identical, NaN where the reference's normalize(sum gradW) is NaN
    assert np.array_equal(np.isnan(nrm), np.isnan(onrm))
    ok = ~np.isnan(onrm)
    assert np.array_equal(nrm[ok], onrm[ok])
    assert stats["samples"] == ocnt["samples"] and stats["hits"] == ocnt["hits"]


def test_reference_mode_pool_skip(marcher, pool):
    marcher.set_frame(10, pool.xyz, pool.h)
    c = pool.cameras[0]
    G = O.Grid(pool.xyz, pool.h)
    depth = depth_in_front(c, G.bmin, G.bmax, 0.02)
    for skip in (2, 3):
        st = dict(mode=1, max_steps=400, step=0.004, iso=pool.iso, normal_mode=1, skip_mode=skip, compensated=0)
        pos, nrm, t, stats = marcher.march(10, st, c.inv_pv, c.eye, c.width, c.height, depth=depth)
        opos, onrm, ot, ocnt = O.march(G, O.make_settings(**st), c.inv_pv, c.eye, c.width, c.height, depth=depth)
        assert np.array_equal(pos, opos) and np.array_equal(t, ot)
        assert stats["samples"] == ocnt["samples"]


@pytest.mark.parametrize("size", [(203, 77), (64, 8), (300, 9)])
def test_reference_mode_ragged_active_tiles(marcher, size):
    """The reference march runs over the tiles with a drawn depth (the tile
    pre-pass writes the others' misses): ragged image sizes, tiles partly
    outside the image and tiles with no depth at all, against the oracle;
    every pixel is counted once, the started rays are the pixels whose depth
    is not 1."""
    W, H = size
    sc = scenes.lattice_cube(2, width=W, height=H)
    c = sc.cameras[1]
    marcher.set_frame(12, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    depth = depth_in_front(c, G.bmin, G.bmax, 0.05)
    assert 0 < (depth != 1.0).sum() < W * H
    st = dict(mode=1, max_steps=128, step=0.009, iso=sc.iso, normal_mode=1, skip_mode=2, compensated=0)
    pos, nrm, t, stats = marcher.march(12, st, c.inv_pv, c.eye, W, H, depth=depth)
    opos, onrm, ot, ocnt = O.march(G, O.make_settings(**st), c.inv_pv, c.eye, W, H, depth=depth)
    assert np.array_equal(pos, opos) and np.array_equal(t, ot)
    assert np.array_equal(np.isnan(nrm), np.isnan(onrm))
    ok = ~np.isnan(onrm)
    assert np.array_equal(nrm[ok], onrm[ok])
    assert stats["rays"] == W * H and stats["rays_in_aabb"] == int((depth != 1.0).sum())
    assert stats["samples"] == ocnt["samples"] and stats["hits"] == ocnt["hits"]
