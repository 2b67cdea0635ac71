"""Input and output files of tests/facade/facade_demo.cpp, the program that
drives the RayMarcher facade as AdvancedRenderer does (Prepare, Start, poll
IsDone; AdvancedRenderer.cpp:257-298)."""
import pathlib
import struct
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
DEMO = ROOT / "bachelor-thesis_amd" / "facade_demo"


def run_demo(tmp_path, sc, c, depth, max_steps, step, iso, aniso, with_log=False):
    """Runs facade_demo on one view; its images (and its stderr with with_log)."""
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
    proc = subprocess.run([str(DEMO), str(inp), str(out)], check=True, timeout=120, capture_output=True, text=True)
    print(proc.stderr, flush=True)
    raw = np.fromfile(out, np.float32)
    pos, nrm = raw[: 4 * W * H].reshape(H, W, 4), raw[4 * W * H: 8 * W * H].reshape(H, W, 4)
    return (pos, nrm, proc.stderr) if with_log else (pos, nrm)
