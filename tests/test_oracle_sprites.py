"""The oracle's depth-sprite pass (DepthRenderPass, depth.vert/frag, the
sprite quads of AdvancedRenderer.cpp:455-474) pinned by what the geometry
fixes: the spherical offset's cosine against numpy; a sprite straight ahead
of the camera, whose centre depth is the projection of the point R in front of
the particle and whose rim depth is that of the particle's plane (FP64 closed
forms of the left-handed [0, 1] perspective); its pixel count against the
projected disc's area; overlapping sprites keep the per-pixel minimum; sprites
behind the camera or beyond the far plane draw nothing."""
import math

import numpy as np
import pytest

import oracle as O
from pmarch.camera import make_camera


def test_cos_half_pi():
    s = np.linspace(0.0, 1.0, 2001, dtype=np.float32)
    c = O.cos_half_pi(s)
    assert np.abs(c - np.cos(np.pi / 2 * s.astype(np.float64))).max() < 3e-7
    assert c[0] == 1.0 and abs(float(c[-1])) < 1e-7


def _ndc_z(cam, zv):
    """[0, 1] depth of view-space z (glm::perspectiveLH_ZO)."""
    n, f = cam.near, cam.far
    return (f / (f - n) * zv - f * n / (f - n)) / zv


def _axis_cam(W=65, H=65, fovy=40.0):
    return make_camera((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), W, H, fovy_deg=fovy, near=0.1, far=100.0)


def test_sprite_on_axis_closed_forms():
    cam = _axis_cam()
    R, zc = 0.1, 2.0
    d = O.depth_sprites(np.array([[0.0, 0.0, zc]], np.float32), cam, R)
    drawn = d < 1.0
    # the centre pixel looks along the axis: the sphere's front point, R nearer
    assert d[32, 32] == pytest.approx(_ndc_z(cam, zc - R), rel=2e-6)
    assert d.min() == d[32, 32]
    # every pixel in FP64: the centre ray meets the plane z = zc at (x, y);
    # l2 = (x^2 + y^2) / R^2 decides coverage, the depth is that of
    # zc - R cos(pi/2 sqrt(l2)) (the rim, l2 -> 1, tends to the plane's depth)
    t = math.tan(math.radians(40.0) / 2)
    ndx = (np.arange(65) + 0.5) * (2.0 / 65) - 1.0
    x, y = ndx[None, :] * t * zc, -ndx[:, None] * t * zc
    l2 = (x ** 2 + y ** 2) / R ** 2
    clear = np.abs(l2 - 1.0) > 1e-4
    assert np.array_equal(drawn[clear], (l2 <= 1.0)[clear])
    want = _ndc_z(cam, zc - R * np.cos(np.pi / 2 * np.sqrt(np.minimum(l2, 1.0))))
    assert np.allclose(d[drawn], want[drawn], rtol=3e-6, atol=0)
    assert d[drawn].max() < _ndc_z(cam, zc)
    # pixels covered: the disc of radius R at distance zc, in pixels
    r_px = R / zc / math.tan(math.radians(40.0) / 2) * 65 / 2
    assert abs(drawn.sum() - math.pi * r_px ** 2) < 2 * math.pi * r_px + 4
    # symmetric about the centre pixel
    assert np.array_equal(drawn, drawn[::-1, :]) and np.array_equal(drawn, drawn[:, ::-1])


def test_sprites_keep_the_minimum():
    cam = make_camera((0.3, 0.4, -1.5), (0.5, 0.5, 0.5), 96, 72, fovy_deg=50.0)
    rng = np.random.default_rng(3)
    a = rng.uniform(0.3, 0.7, (40, 3)).astype(np.float32)
    b = rng.uniform(0.3, 0.7, (40, 3)).astype(np.float32)
    R = 0.05
    da, db = O.depth_sprites(a, cam, R), O.depth_sprites(b, cam, R)
    dab = O.depth_sprites(np.concatenate([a, b]), cam, R)
    assert np.array_equal(dab, np.minimum(da, db))
    assert np.array_equal(dab, O.depth_sprites(np.concatenate([b, a])[::-1], cam, R))
    assert (dab < 1.0).sum() > 150 and ((da < 1.0) & (db < 1.0)).sum() > 10    # overlaps occur


def test_sprites_behind_or_beyond_draw_nothing():
    cam = _axis_cam()
    behind = np.array([[0.0, 0.0, -2.0], [0.0, 0.0, 0.05]], np.float32)       # behind; straddling the eye
    assert (O.depth_sprites(behind, cam, 0.1) == 1.0).all()
    far = np.array([[0.0, 0.0, 150.0]], np.float32)                            # past the far plane (100)
    assert (O.depth_sprites(far, cam, 1.0) == 1.0).all()
    assert (O.depth_sprites(np.zeros((0, 3), np.float32), cam, 0.1) == 1.0).all()
