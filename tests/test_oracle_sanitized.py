"""The CPU oracle (test infrastructure, plain C) built with AddressSanitizer
and UBSan and driven through its march modes (configured and reference-style
hashed, every skip, both normal forms, isotropic and anisotropic, FP64),
the density queries, the extended walk, the depth smoothing and the depth
sprites, in a child
process with the sanitizer runtimes preloaded: no memory or undefined-
behaviour report.  No GPU."""
import os
import pathlib
import subprocess
import sys
import textwrap

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent

SCRIPT = textwrap.dedent("""
    import pathlib, sys
    sys.path[:0] = [{pkg!r}, {orc!r}]
    import oracle as O
    O.LIB_PATH = pathlib.Path({lib!r})
    import numpy as np
    from pmarch import scenes
    sc = scenes.lattice_cube(0, width=96, height=72)
    G = O.Grid(sc.xyz, sc.h)
    cam = sc.cameras[0]
    for st in (sc.settings(), sc.settings(skip_mode=4), sc.settings(skip_mode=1, normal_mode=1),
               scenes.aniso_settings(sc, skip_mode=4), scenes.aniso_settings(sc, normal_mode=1)):
        O.march(G, O.make_settings(**st), cam.inv_pv, cam.eye, cam.width, cam.height, threads=4)
    O.march64(G, O.make_settings(**sc.settings()), cam.inv_pv, cam.eye, cam.width, cam.height, threads=4)
    q = sc.xyz[:50] + 0.01
    G.density(q); G.density64(q); G.density_aniso(O.make_settings(**scenes.aniso_settings(sc)), q)
    G.neighbors_ext(q[0], 0.2); G.cdiff_union(q[0], sc.normal_eps)
    O.depth_smooth(np.random.default_rng(0).uniform(0, 1, (20, 30)).astype(np.float32), 8)
    O.depth_sprites(sc.xyz, cam, sc.h)
    from pmarch.camera import make_camera
    inside = make_camera(sc.xyz.mean(0), sc.xyz.mean(0) + [0.0, 0.0, 1.0], 31, 17)
    O.depth_sprites(sc.xyz, inside, 4 * sc.h)           # quads behind the eye, quads past the image
    O.HashedSearch(sc.xyz, sc.h).march(G, O.make_settings(**sc.settings()), cam.inv_pv, cam.eye, 96, 72, threads=4)
    print("done")
""")


def test_oracle_under_asan_ubsan(tmp_path):
    try:
        asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
        ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    except FileNotFoundError:
        pytest.skip("gcc not available")
    if not (os.path.isabs(asan) and os.path.exists(asan) and os.path.isabs(ubsan) and os.path.exists(ubsan)):
        pytest.skip("sanitizer runtimes not available")
    lib = tmp_path / "liboracle_asan.so"
    subprocess.run(["gcc", "-O1", "-g", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-shared",
                    "-pthread", "-o", str(lib), str(ROOT / "oracle" / "pm_oracle.c"), "-lm"], check=True)
    script = tmp_path / "run.py"
    script.write_text(SCRIPT.format(pkg=str(ROOT / "bachelor-thesis_amd"), orc=str(ROOT / "oracle"), lib=str(lib)))
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("done"), r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
