"""The RayMarcher facade (bachelor-thesis_amd/csrc/facade) driven like
AdvancedRenderer drives the reference (Prepare, Start, poll IsDone;
AdvancedRenderer.cpp:257-298): its images equal the C ABI's reference march and
the CPU oracle's."""
import numpy as np
import pytest

import oracle as O
from facade_io import DEMO, run_demo as _run_demo
from parity import depth_in_front
from pmarch import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("aniso", [0, 1])
def test_facade_reference_march(marcher, tmp_path, aniso):
    if not DEMO.exists():
        pytest.fail("facade_demo not built (run __graft_entry__.build())")
    sc = scenes.lattice_cube(0, width=200, height=150)
    c = sc.cameras[0]
    G = O.Grid(sc.xyz, sc.h)
    depth = depth_in_front(c, G.bmin, G.bmax, 0.05)
    W, H = c.width, c.height
    iso = np.float32(1.0) if aniso else np.float32(sc.iso)
    step, max_steps = np.float32(0.009), 128
    pos, nrm = _run_demo(tmp_path, sc, c, depth, max_steps, step, iso, aniso)
    st = dict(mode=1, max_steps=max_steps, step=float(step), iso=float(iso), normal_mode=1, skip_mode=2, compensated=0)
    if aniso:
        st.update(scenes.ANISO_SETTINGS)
    marcher.set_frame(5, sc.xyz, sc.h)
    gp, gn, _, stats = marcher.march(5, st, c.inv_pv, c.eye, W, H, depth=depth)
    assert np.array_equal(pos, gp)
    assert np.array_equal(np.isnan(nrm), np.isnan(gn)) and np.array_equal(nrm[~np.isnan(nrm)], gn[~np.isnan(gn)])
    op, on, _, _ = O.march(G, O.make_settings(**st), c.inv_pv, c.eye, W, H, depth=depth)
    assert np.array_equal(pos, op)
    assert (pos[..., 3] > 0).sum() > 5000 and stats["hits"] == int((pos[..., 3] > 0).sum())
