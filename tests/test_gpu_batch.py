"""Batched views x frames in one persistent launch (plain and learned), with
and without hipGraph capture: every view equals the single-view march of the
same frame and camera and the oracle's march, bit for bit, and graph replays
equal the eager run."""
import numpy as np
import pytest

import oracle as O
import pmarch
from pmarch import scenes, train

pytestmark = pytest.mark.gpu


def _frames(marcher, base_id=60, n=3, with_xyz=False):
    sc = scenes.dam_break(7, nx=30, ny=60, nz=30)
    frames = scenes.displaced_frames(sc, n, seed=3, spacing=0.005)
    for i, xyz in enumerate(frames):
        marcher.set_frame(base_id + i, xyz, sc.h)
    ids = [base_id + i for i in range(n)]
    if with_xyz:
        return sc, ids, {base_id + i: xyz for i, xyz in enumerate(frames)}
    return sc, ids


def _views(sc, fids, n, seed):
    rng = np.random.default_rng(seed)
    cams = scenes.random_views(rng, (0.075, 0.15, 0.075), n, 40, 300, 30, 200, radius=(0.35, 0.6))
    return [(fids[i % len(fids)], c.width, c.height, c.inv_pv, c.eye) for i, c in enumerate(cams)]


def _check_views(marcher, sc, views, settings, frames=None):
    """Every view against the single-view march; with the frames' positions,
    also against the oracle's (plain) march of the same frame and camera."""
    keys = dict(O.Settings._fields_)
    osettings = O.make_settings(**{k: v for k, v in sc.settings().items() if k in keys})
    grids = {}
    for i, (f, W, H, ipv, eye) in enumerate(views):
        b = marcher.batch_download(i)
        s = marcher.march(f, settings, ipv, eye, W, H)
        for x, y in zip(b, s[:3]):
            assert np.array_equal(x, y), f"view {i} ({W}x{H}, frame {f}) differs"
        if frames is not None:
            if f not in grids:
                grids[f] = O.Grid(frames[f], sc.h)
            o = O.march(grids[f], osettings, ipv, eye, W, H)
            for x, y in zip(b, o[:3]):
                assert np.array_equal(x, y), f"view {i} ({W}x{H}, frame {f}) differs from the oracle"


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_plain_matches_single_views(marcher, use_graph):
    sc, fids, xyz = _frames(marcher, with_xyz=True)
    views = _views(sc, fids, 12, seed=1)
    st = sc.settings(skip_mode=4)
    print(f"batch plain graph={use_graph} views={len(views)}", flush=True)
    marcher.batch_set(st, views, use_graph=use_graph)
    marcher.batch_run()
    marcher.wait()
    bs = marcher.stats()
    assert bs["rays"] == sum(W * H for _, W, H, _, _ in views)
    _check_views(marcher, sc, views, st, frames=xyz)


def test_batch_graph_replays_match_eager(marcher):
    sc, fids = _frames(marcher)
    views = _views(sc, fids, 8, seed=2)
    st = sc.settings(skip_mode=4)
    print("batch graph replays", flush=True)
    marcher.batch_set(st, views, use_graph=False)
    marcher.batch_run()
    marcher.wait()
    eager = [marcher.batch_download(i) for i in range(len(views))]
    eager_stats = marcher.stats()
    marcher.batch_set(st, views, use_graph=True)
    for rep in range(3):
        marcher.batch_run()
        marcher.wait()
        s = marcher.stats()
        for k in ("rays", "hits", "samples", "march_cand", "march_samples"):
            assert s[k] == eager_stats[k], (rep, k)
        for i in range(len(views)):
            for x, y in zip(marcher.batch_download(i), eager[i]):
                assert np.array_equal(x, y), f"replay {rep} view {i}"
        # eager work between replays on the same stream
        f, W, H, ipv, eye = views[rep]
        marcher.march(f, st, ipv, eye, W, H)


def test_batch_learned_matches_single_views(marcher):
    if not train.WEIGHTS.exists():
        pytest.skip("no trained network checkpoint")
    params, _ = train.load()
    marcher.mlp_set(params)
    sc, fids, xyz = _frames(marcher, with_xyz=True)
    views = _views(sc, fids, 10, seed=4)
    st = sc.settings(learned=1)
    print("batch learned", flush=True)
    marcher.batch_set(st, views, use_graph=True)
    marcher.batch_run()
    marcher.wait()
    assert marcher.stats()["mlp_evals"] > 0
    _check_views(marcher, sc, views, st, frames=xyz)


def test_batch_invalidated_by_frame_update(marcher):
    sc, fids = _frames(marcher, base_id=70, n=2)
    views = _views(sc, fids, 3, seed=5)
    marcher.batch_set(sc.settings(skip_mode=4), views, use_graph=True)
    marcher.set_frame(fids[0], sc.xyz, sc.h)
    with pytest.raises(pmarch.PmError):
        marcher.batch_run()


def test_batch_rejects_bad_views(marcher):
    sc, fids = _frames(marcher, base_id=72, n=1)
    with pytest.raises(pmarch.PmError):
        marcher.batch_set(sc.settings(), [(999, 64, 64, np.eye(4), np.zeros(3))])
    with pytest.raises(pmarch.PmError):
        marcher.batch_set(sc.settings(), [(fids[0], 0, 64, np.eye(4), np.zeros(3))])
    with pytest.raises(pmarch.PmError):
        marcher.batch_set(sc.settings(mode=1), [(fids[0], 64, 64, np.eye(4), np.zeros(3))])
