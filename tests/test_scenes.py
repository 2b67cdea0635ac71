"""Workload generators: sizes, seeding and the iso/step conventions -- no GPU."""
import math

import numpy as np
import pytest

from pmarch import camera as pc
from pmarch import scenes


def test_lattice_cube_shape_and_seed():
    a = scenes.lattice_cube(0)
    b = scenes.lattice_cube(0)
    c = scenes.lattice_cube(1)
    assert a.n == 32 ** 3
    assert np.array_equal(a.xyz, b.xyz) and not np.array_equal(a.xyz, c.xyz)
    assert a.h == pytest.approx(0.1) and a.step == pytest.approx(0.0125) and a.normal_eps == pytest.approx(0.003125)
    span = a.xyz.max(0) - a.xyz.min(0)
    assert np.all(span > 1.5) and np.all(span < 1.56)          # overhangs the unit box
    assert a.iso == pytest.approx(0.5 * 8 / (math.pi * 0.1 ** 3) * 3.1415061532636885, rel=1e-6)
    cam = a.cameras[0]
    assert np.linalg.norm(cam.eye - 0.5) == pytest.approx(2.0, rel=1e-6)
    assert (cam.width, cam.height) == (640, 480)


def test_dam_break_counts():
    s = scenes.dam_break(0)
    assert s.n == 1_024_000
    assert s.h == pytest.approx(0.01) and s.step == pytest.approx(0.00125)
    cam = s.cameras[0]
    assert (cam.width, cam.height) == (1920, 1080)
    # orbit camera at 30 degrees elevation
    v = cam.eye - np.array(scenes.COLUMN_CENTRE)
    assert math.degrees(math.asin(v[1] / np.linalg.norm(v))) == pytest.approx(30.0, abs=1e-4)
    jit = s.xyz.astype(np.float64) - (np.floor(s.xyz / 0.005) + 0.5) * 0.005
    assert np.std(s.xyz[:, 0] - np.round((s.xyz[:, 0] - 0.0025) / 0.005) * 0.005 - 0.0025) < 0.003
    assert jit.shape == s.xyz.shape


def test_displaced_frames_bounded():
    base = scenes.dam_break(0, nx=8, ny=16, nz=8)
    fr = scenes.displaced_frames(base, 4, seed=3)
    assert len(fr) == 4
    for a, b in zip(fr, fr[1:]):
        assert np.linalg.norm(b.astype(np.float64) - a, axis=1).max() <= 0.01 * 0.005 * 1.0001


def test_camera_projects_target_to_centre():
    cam = pc.make_camera((0.3, 1.1, -1.7), (0.2, 0.4, 0.2), 320, 200)
    ph = cam.pv @ np.array([0.2, 0.4, 0.2, 1.0])
    assert abs(ph[0] / ph[3]) < 1e-12 and abs(ph[1] / ph[3]) < 1e-12
    assert 0.0 < ph[2] / ph[3] < 1.0


def test_clustered_pool_small():
    s = scenes.clustered_pool(0, n_total=100_000, width=64, height=36)
    assert s.n == 100_000
    assert s.h == pytest.approx(0.012)
    assert np.sum(s.xyz[:, 1] <= 0.15 + 1e-3) >= 60_000
