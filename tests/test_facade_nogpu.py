"""The RayMarcher facade without a GPU: the reference contract (RayMarcher.h:
28-77; ThreadPool.cpp:33-36 -- Start returns, IsDone turns true, the caller
uploads the images) still holds, and the images it hands back are misses
(exact zeros) rather than whatever the caller's staging buffers held; the
reason is reported.  Runs where no GPU is visible (the CPU suite)."""
import ctypes

import numpy as np
import pytest

import pmarch
from facade_io import DEMO, run_demo
from pmarch import scenes


def test_facade_without_gpu_writes_misses(tmp_path):
    L = pmarch.load()
    p = ctypes.c_void_p()
    if L.pm_create(0, ctypes.byref(p)) == 0:
        L.pm_destroy(p)
        pytest.skip("a GPU is present: the facade's GPU path is tested in test_gpu_facade.py")
    if not DEMO.exists():
        pytest.fail("facade_demo not built (run __graft_entry__.build())")
    sc = scenes.lattice_cube(0, width=40, height=30)
    c = sc.cameras[0]
    depth = np.full((c.height, c.width), 0.5, np.float32)
    pos, nrm, log = run_demo(tmp_path, sc, c, depth, 64, 0.009, sc.iso, 0, with_log=True)
    # facade_demo fills both buffers with -7 before Prepare
    assert np.all(pos == 0.0) and np.all(nrm == 0.0)
    assert "no GPU" in log
