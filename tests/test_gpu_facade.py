"""The RayMarcher facade (bachelor-thesis_amd/csrc/facade) driven like
AdvancedRenderer drives the reference (Prepare, Start, poll IsDone;
AdvancedRenderer.cpp:257-298): its images equal the C ABI's reference march and
the CPU oracle's."""
import pathlib
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from parity import depth_in_front
from pmarch import scenes

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parent.parent
DEMO = ROOT / "bachelor-thesis_amd" / "facade_demo"


def _run_demo(tmp_path, sc, c, depth, max_steps, step, iso, aniso):
    W, H = c.width, c.height
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(b"PMFD")
        f.write(struct.pack("<3I", W, H, sc.n))
        f.write(np.float32(sc.h).tobytes())
        f.write(np.ascontiguousarray(c.inv_pv, np.float32).tobytes())
        f.write(np.ascontiguousarray(c.eye, np.float32).tobytes())
        f.write(struct.pack("<i", max_steps))
        f.write(np.float32(step).tobytes())
        f.write(np.float32(iso).tobytes())
        # EnableAnisotropy, k_n, k_r, k_s, N_eps (AdvancedRenderer.cpp:18-28), ParticleRadiusExt
        f.write(struct.pack("<i3fi", int(aniso), 0.5, 2.0, 2000.0, 1))
        f.write((np.float32(2.0) * np.float32(sc.h)).tobytes())
        f.write(np.ascontiguousarray(depth, np.float32).tobytes())
        f.write(np.ascontiguousarray(sc.xyz, np.float32).tobytes())
    out = tmp_path / "out.bin"
    print(f"facade_demo {W}x{H} anisotropic={aniso}", flush=True)
    subprocess.run([str(DEMO), str(inp), str(out)], check=True, timeout=120)
    raw = np.fromfile(out, np.float32)
    return raw[: 4 * W * H].reshape(H, W, 4), raw[4 * W * H: 8 * W * H].reshape(H, W, 4)


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
