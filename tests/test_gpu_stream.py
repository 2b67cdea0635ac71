"""Dataset IO and frame streaming on the GPU: a frame loaded from a BGEO file
(Dataset::ReadFile, Dataset.cpp:292-306) builds the same grids and images as
the same positions set directly; a frame built on the second stream while a
march of another frame is in flight (the reference's autoplay,
AdvancedRenderer.cpp:283-290) leaves that march's images unchanged and then
marches exactly like a frame set synchronously, and like the oracle."""
import time

import numpy as np
import pytest

import oracle as O
from pmarch import PmError, scenes
from test_bgeo import write_bgeo

pytestmark = pytest.mark.gpu


def _same_grid(m, a, b):
    fa, csa, sa, ia = m.grid(a)
    fb, csb, sb, ib = m.grid(b)
    assert tuple(fa.dims) == tuple(fb.dims) and tuple(fa.origin) == tuple(fb.origin)
    assert np.array_equal(csa, csb) and np.array_equal(sa, sb) and np.array_equal(ia, ib)


def test_load_bgeo_equals_set_frame(marcher, tmp_path):
    sc = scenes.lattice_cube(1)
    cam = sc.cameras[0]
    f = tmp_path / "lattice_1.bgeo.gz"
    write_bgeo(f, sc.xyz, [("v", 3, 5, None), ("id", 1, 1, None)], gz=True)
    print("load_bgeo lattice cube", flush=True)
    marcher.load_bgeo(4, f, sc.h)
    marcher.set_frame(5, sc.xyz, sc.h)
    _same_grid(marcher, 4, 5)
    st = sc.settings(skip_mode=4)
    a = marcher.march(4, st, cam.inv_pv, cam.eye, cam.width, cam.height)
    b = marcher.march(5, st, cam.inv_pv, cam.eye, cam.width, cam.height)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def test_stream_frame_beside_march(marcher):
    base = scenes.dam_break(0)
    frames = scenes.displaced_frames(base, 3, seed=3)
    cam = base.cameras[0]
    st = base.settings(skip_mode=4)
    print("stream: dam break frames 0..2", flush=True)
    marcher.set_frame(0, frames[0], base.h)
    ref0 = marcher.march(0, st, cam.inv_pv, cam.eye, cam.width, cam.height)
    # march frame 0; build frame 1 beside it on the second stream
    pos, nrm, t = marcher.march_async(0, st, cam.inv_pv, cam.eye, cam.width, cam.height)
    t0 = time.perf_counter()
    marcher.set_frame_stream(1, frames[1], base.h)
    t_build = time.perf_counter() - t0
    with pytest.raises(PmError):
        marcher.set_frame_stream(0, frames[1], base.h)        # the march in flight reads frame 0
    marcher.wait()
    for x, y in zip((pos, nrm, t), ref0[:3]):
        assert np.array_equal(x, y)
    # frame 1 built on the second stream marches like frame 1 set synchronously
    got = marcher.march(1, st, cam.inv_pv, cam.eye, cam.width, cam.height)
    marcher.set_frame(2, frames[1], base.h)
    _same_grid(marcher, 1, 2)
    want = marcher.march(2, st, cam.inv_pv, cam.eye, cam.width, cam.height)
    for x, y in zip(got[:3], want[:3]):
        assert np.array_equal(x, y)
    # and equals the oracle's march of frame 1 (every 27th row)
    step = 27
    keys = dict(O.Settings._fields_)
    op, on, ot, _ = O.march(O.Grid(frames[1], base.h), O.make_settings(**{k: v for k, v in st.items() if k in keys}),
                            cam.inv_pv, cam.eye, cam.width, cam.height, row_step=step)
    for x, y in zip(got[:3], (op, on, ot)):
        assert np.array_equal(x[::step], y[::step])
    print(f"frame build beside the march: {t_build * 1e3:.2f} ms host time", flush=True)


def test_autoplay_double_buffer(marcher):
    """Frames 0..3 through two frame slots: march slot i % 2 while the next
    frame streams into the other (with the bound skip's bits for the last
    march's iso built beside it too); every image and counter equals the
    frame's own march, and a march with another iso (bits rebuilt) as well."""
    base = scenes.dam_break(1, nx=40, ny=80, nz=40, width=640, height=360)
    frames = scenes.displaced_frames(base, 4, seed=4)
    cam = base.cameras[0]
    st = base.settings(skip_mode=4)
    print("autoplay: 4 frames, 2 slots", flush=True)
    marcher.set_frame(10, frames[0], base.h)
    imgs = []
    for i in range(4):
        slot = 10 + (i % 2)
        pos, nrm, t = marcher.march_async(slot, st, cam.inv_pv, cam.eye, cam.width, cam.height)
        if i + 1 < 4:
            marcher.set_frame_stream(10 + ((i + 1) % 2), frames[i + 1], base.h)
        marcher.wait()
        imgs.append((pos, nrm, t, marcher.stats()))
    keys = ("rays", "rays_in_aabb", "hits", "samples", "march_cand", "march_samples")
    for i in range(4):
        marcher.set_frame(12, frames[i], base.h)
        want = marcher.march(12, st, cam.inv_pv, cam.eye, cam.width, cam.height)
        for x, y in zip(imgs[i][:3], want[:3]):
            assert np.array_equal(x, y), i
        assert all(imgs[i][3][k] == want[3][k] for k in keys), i
    # slot 11 holds frame 3 with the bits built for st's iso; another iso rebuilds them
    st2 = base.settings(skip_mode=4, iso=base.iso * 1.5)
    got = marcher.march(11, st2, cam.inv_pv, cam.eye, cam.width, cam.height)
    want = marcher.march(12, st2, cam.inv_pv, cam.eye, cam.width, cam.height)
    for x, y in zip(got[:3], want[:3]):
        assert np.array_equal(x, y)
    assert all(got[3][k] == want[3][k] for k in keys)
