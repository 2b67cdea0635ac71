"""The learned step at the configured sizes, with the committed network
(pmarch/weights/step_mlp.npz: 2 M rows of plain-march traces over 20 dam-break
and 20 clustered-pool scenes).

* configs[3] scenes: the 1,024,000-particle dam break at 1920x1080 and the
  4,000,000-particle clustered pool at 3840x2160, learned step on: the whole
  image equals the plain march's bit for bit (every skipped sample is
  certified below iso), row samples equal the CPU oracle's plain march, and
  the network did work (proposals, certified skips).
* 64 views (W ~ U{256..1920}, H ~ U{256..1080}) x 16 frames of the
  1 M dam break (1 % displacement per frame) in one persistent launch replayed
  from a hipGraph, learned step on: row samples of a spread of views equal the
  oracle's march of the same frame and camera, and every one of the 1,024
  view-frames is written once (misses exactly zero, hits with w = 1, unit
  normals) with its hit points on their rays at the reported t.
"""
import numpy as np
import pytest

import oracle as O
from pmarch import scenes, train

pytestmark = pytest.mark.gpu

KEYS = dict(O.Settings._fields_)


def _oracle_settings(sc):
    return O.make_settings(**{k: v for k, v in sc.settings().items() if k in KEYS})


def _rows_equal(gpu, ora, step):
    for a, b in zip(gpu, ora):
        assert np.array_equal(a[::step], b[::step])


@pytest.fixture(scope="module")
def network(marcher):
    params, summary = train.load()
    print("network", {k: summary[k] for k in ("rows", "scenes", "epochs")}, flush=True)
    return params


@pytest.mark.parametrize("boxes", [0, 1])
@pytest.mark.parametrize("name", ["dam_break_1080p", "clustered_pool_4k"])
def test_learned_configured_scenes(marcher, network, name, boxes):
    """With the 4x4x4 bound (boxes = 1) the network proposes and certified jumps
    are taken; with 8x8x8 (boxes = 0) the bound itself passes almost every
    low-density sample, so few proposals remain.  The image is the plain
    march's either way."""
    if name == "dam_break_1080p":
        sc, step = scenes.dam_break(0), 27
    else:
        sc, step = scenes.clustered_pool(0), 72
    c = sc.cameras[0]
    print(f"learned {name} n={sc.n} {c.width}x{c.height} bound_boxes={boxes}", flush=True)
    marcher.set_frame(70, sc.xyz, sc.h)
    marcher.mlp_set(network)
    plain = marcher.march(70, sc.settings(skip_mode=4, bound_boxes=boxes), c.inv_pv, c.eye, c.width, c.height)
    lrn = marcher.march(70, sc.settings(skip_mode=4, bound_boxes=boxes, learned=1), c.inv_pv, c.eye, c.width,
                        c.height)
    sp, sl = plain[3], lrn[3]
    print("plain", sp, "\nlearned", sl, flush=True)
    for a, b in zip(plain[:3], lrn[:3]):
        assert np.array_equal(a, b)
    assert sl["march_samples"] <= sp["march_samples"]
    if boxes == 1:
        assert sl["mlp_evals"] > 1000 and sl["cert_skipped"] > 1000
        assert sl["march_samples"] < sp["march_samples"]
    op, on, ot, _ = O.march(O.Grid(sc.xyz, sc.h), _oracle_settings(sc), c.inv_pv, c.eye, c.width, c.height,
                            row_step=step)
    _rows_equal(lrn[:3], (op, on, ot), step)


def _written_once(pos, nrm, t, eye, inv_dir):
    """Every pixel is a miss (all zero, t = -1) or a hit (w = 1, unit normal,
    the position on the pixel's ray at distance t)."""
    hit = pos[..., 3] > 0
    assert np.all(pos[~hit] == 0) and np.all(nrm[~hit] == 0) and np.all(t[~hit] == -1.0)
    assert np.all(pos[hit][:, 3] == 1.0) and np.all(nrm[hit][:, 3] == 1.0)
    n = np.linalg.norm(nrm[hit][:, :3].astype(np.float64), axis=1)
    assert np.all(np.abs(n - 1.0) < 1e-5)
    p = pos[hit][:, :3].astype(np.float64)
    v = p - eye.astype(np.float64)
    dist = np.linalg.norm(v, axis=1)
    assert np.all(np.abs(dist - t[hit]) <= 1e-5 * np.maximum(1.0, dist))
    d = inv_dir[hit]
    cosang = (v * d).sum(1) / np.maximum(dist, 1e-30)
    assert np.all(cosang > 1.0 - 1e-9)
    return int(hit.sum())


def _ray_dirs(cam):
    """Unit directions of the pixel-corner rays (FP64)."""
    W, H = cam.width, cam.height
    M = cam.inv_pv.reshape(4, 4).T.astype(np.float64)
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    cx, cy = xs * 2.0 / W - 1.0, ys * 2.0 / H - 1.0
    ph = np.einsum("ij,jhw->hwi", M, np.stack([cx, cy, np.ones_like(cx), np.ones_like(cx)]))
    d = ph[..., :3] / ph[..., 3:] - cam.eye.astype(np.float64)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_learned_batch_64_views_16_frames(marcher, network):
    base = scenes.dam_break(0)
    frames = scenes.displaced_frames(base, 16, seed=1)
    for f, xyz in enumerate(frames):
        marcher.set_frame(100 + f, xyz, base.h)
    rng = np.random.default_rng(2024)
    cams = scenes.random_views(rng, scenes.COLUMN_CENTRE, 64, 256, 1920, 256, 1080, radius=(1.1, 1.9))
    views = [(100 + f, c.width, c.height, c.inv_pv, c.eye) for c in cams for f in range(16)]
    rays = sum(W * H for _, W, H, _, _ in views)
    print(f"learned batch: {len(views)} view-frames, {rays} rays", flush=True)
    marcher.mlp_set(network)
    st = base.settings(skip_mode=4, learned=1)
    marcher.batch_set(st, views, use_graph=True)
    for rep in range(2):
        marcher.batch_run()
        marcher.wait()
    s = marcher.stats()
    print(s, flush=True)
    assert s["rays"] == rays and s["mlp_evals"] > 0 and s["cert_skipped"] > 0
    # every view-frame: written once, hit points on their rays
    hits = 0
    dirs = {}
    for i, (f, W, H, ipv, eye) in enumerate(views):
        ci = i // 16
        if ci not in dirs:
            dirs = {ci: _ray_dirs(cams[ci])}
        pos, nrm, t = marcher.batch_download(i)
        hits += _written_once(pos, nrm, t, eye, dirs[ci])
    assert hits == s["hits"]
    # a spread of views (all 16 frames, 8 cameras) against the oracle's march
    grids = {}
    for ci in range(0, 64, 8):
        f = (ci // 8) * 2 + 1
        i = ci * 16 + f
        _, W, H, ipv, eye = views[i]
        if f not in grids:
            grids = {f: O.Grid(frames[f], base.h)}
        step = max(1, H // 16)
        op, on, ot, _ = O.march(grids[f], _oracle_settings(base), ipv, eye, W, H, row_step=step)
        print(f"view {i}: camera {ci} frame {f} {W}x{H} vs oracle every {step}th row", flush=True)
        _rows_equal(marcher.batch_download(i), (op, on, ot), step)
