"""Full-size workloads on the GPU: the 4 M-particle clustered pool at 3840x2160
(bit-exact against the oracle on a row sample, and through size-independent
properties on the whole image), the 1080p dam break against the oracle on a
row sample, and every form's error against the FP64 oracle march."""
import numpy as np
import pytest

import oracle as O
from parity import assert_within, compare, extent
from pmarch import scenes

pytestmark = pytest.mark.gpu

KEYS = dict(O.Settings._fields_)


def _rows_equal(gpu, ora, row_step):
    for a, b in zip(gpu, ora):
        assert np.array_equal(a[::row_step], b[::row_step])


@pytest.fixture(scope="module")
def pool4k():
    return scenes.clustered_pool(0)


def test_clustered_pool_4k_rows_match_oracle(marcher, pool4k):
    sc = pool4k
    c = sc.cameras[0]
    print(f"clustered_pool n={sc.n} {c.width}x{c.height}", flush=True)
    marcher.set_frame(50, sc.xyz, sc.h)
    st = sc.settings(skip_mode=4)
    pos, nrm, t, stats = marcher.march(50, st, c.inv_pv, c.eye, c.width, c.height)
    print(stats, flush=True)
    assert stats["hits"] > 1_000_000
    G = O.Grid(sc.xyz, sc.h)
    step = 72
    op, on, ot, _ = O.march(G, O.make_settings(**{k: v for k, v in sc.settings().items() if k in KEYS}), c.inv_pv,
                            c.eye, c.width, c.height, row_step=step)
    _rows_equal((pos, nrm, t), (op, on, ot), step)


def test_clustered_pool_4k_size_independent(marcher, pool4k):
    """Whole image: the skip and the kernel schedule change no pixel; every
    pixel is written once (misses exactly zero, hits with w = 1); hit points
    lie on the ray at the reported t."""
    sc = pool4k
    c = sc.cameras[0]
    marcher.set_frame(50, sc.xyz, sc.h)
    a = marcher.march(50, sc.settings(skip_mode=0), c.inv_pv, c.eye, c.width, c.height)
    b = marcher.march(50, sc.settings(skip_mode=4), c.inv_pv, c.eye, c.width, c.height)
    d = marcher.march(50, sc.settings(skip_mode=4, schedule=1), c.inv_pv, c.eye, c.width, c.height)
    for x, y, z in zip(a[:3], b[:3], d[:3]):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    pos, nrm, t = a[:3]
    hit = pos[..., 3] > 0
    assert np.all(pos[~hit] == 0) and np.all(nrm[~hit] == 0)
    assert np.all(pos[hit][:, 3] == 1.0) and np.all(nrm[hit][:, 3] == 1.0)
    n = np.linalg.norm(nrm[hit][:, :3].astype(np.float64), axis=1)
    assert np.all(np.abs(n - 1.0) < 1e-5)
    # |pos - cam| equals t along the unit ray direction
    dist = np.linalg.norm(pos[hit][:, :3].astype(np.float64) - np.asarray(c.eye, np.float64), axis=1)
    assert np.max(np.abs(dist - t[hit])) < 1e-5 * max(1.0, float(t[hit].max()))
    print("samples no skip", a[3]["samples"], "bound skip", b[3]["samples"], flush=True)


def test_dam_break_1080p_rows_match_oracle(marcher):
    sc = scenes.dam_break(0)
    c = sc.cameras[0]
    marcher.set_frame(51, sc.xyz, sc.h)
    pos, nrm, t, _ = marcher.march(51, sc.settings(skip_mode=4), c.inv_pv, c.eye, c.width, c.height)
    G = O.Grid(sc.xyz, sc.h)
    step = 27
    op, on, ot, _ = O.march(G, O.make_settings(**{k: v for k, v in sc.settings().items() if k in KEYS}), c.inv_pv,
                            c.eye, c.width, c.height, row_step=step)
    _rows_equal((pos, nrm, t), (op, on, ot), step)


@pytest.mark.parametrize("name", ["dam_break", "clustered"])
def test_error_vs_fp64_oracle(marcher, name):
    """The FP32 march against the FP64 oracle march (every operation in double,
    all bisection steps in double) on a row sample of the full-size view."""
    if name == "dam_break":
        sc = scenes.dam_break(0)
        step = 30
    else:
        sc = scenes.clustered_pool(1, n_total=1_000_000, width=1920, height=1080)
        step = 30
    c = sc.cameras[0]
    marcher.set_frame(52, sc.xyz, sc.h)
    pos, nrm, t, _ = marcher.march(52, sc.settings(skip_mode=4), c.inv_pv, c.eye, c.width, c.height)
    G = O.Grid(sc.xyz, sc.h)
    p64, n64, t64 = O.march64(G, O.make_settings(**{k: v for k, v in sc.settings().items() if k in KEYS}), c.inv_pv,
                              c.eye, c.width, c.height, row_step=step)
    m = compare(pos[::step], nrm[::step], t[::step], p64[::step], n64[::step], t64[::step])
    print(name, "vs FP64:", m, "extent", extent(sc.xyz), flush=True)
    assert_within(m, extent(sc.xyz))


def test_camera_inside_aabb_many_tiles(marcher):
    """A 5120x2880 view from inside the particles' AABB: every one of the
    230,400 tiles is active and most rays cross empty space quickly, so waves of
    the persistent march take far more than the 32 tiles they finish
    themselves; the rest go to the finish kernel.  The image equals the
    one-ray-per-lane schedule and the oracle on a row sample."""
    rng = np.random.default_rng(21)
    blobs = [rng.normal(cen, 0.03, (5000, 3)) for cen in ((0.3, 0.5, 0.7), (0.6, 0.4, 0.8), (0.5, 0.6, 0.5))]
    xyz = np.concatenate(blobs + [rng.uniform(0.0, 1.0, (5000, 3))]).astype(np.float32)
    sc = scenes.tiny_scene(xyz, 0.02, eye=(0.5, 0.5, 0.05), target=(0.5, 0.5, 1.0), width=5120, height=2880)
    c = sc.cameras[0]
    print(f"tile overflow n={sc.n} {c.width}x{c.height}", flush=True)
    marcher.set_frame(80, sc.xyz, sc.h)
    st = sc.settings(skip_mode=4)
    a = marcher.march(80, st, c.inv_pv, c.eye, c.width, c.height)
    b = marcher.march(80, dict(st, schedule=1), c.inv_pv, c.eye, c.width, c.height)
    print(a[3], flush=True)
    assert a[3]["rays_in_aabb"] == c.width * c.height
    assert a[3]["hits"] > 100_000
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    for k in ("hits", "samples", "march_cand", "march_samples"):
        assert a[3][k] == b[3][k], k
    G = O.Grid(sc.xyz, sc.h)
    step = 360
    op, on, ot, _ = O.march(G, O.make_settings(**{k: v for k, v in st.items() if k in KEYS}), c.inv_pv,
                            c.eye, c.width, c.height, row_step=step)
    _rows_equal(a[:3], (op, on, ot), step)
