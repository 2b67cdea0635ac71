"""Camera matrices with the reference's conventions.

The reference uses GLM left-handed with depth in [0, 1]
(src/engine/hzpch.h:22-24), a y-flipped perspective projection
(src/engine/camera/Camera3D.cpp:11) and hands the marcher
``InvProjectionView`` and the camera position (RayMarcher.cpp:95-96).  Matrices
are built in float64 and rounded once to float32 on the caller's side, as
``glm::inverse`` would hand them over; they are column-major, 16 floats.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def perspective_lh_zo(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    """glm::perspectiveLH_ZO as a row-major 4x4 (M[row, col])."""
    t = math.tan(fovy / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / (aspect * t)
    m[1, 1] = 1.0 / t
    m[2, 2] = far / (far - near)
    m[3, 2] = 1.0
    m[2, 3] = -(far * near) / (far - near)
    return m


def look_at_lh(eye, center, up) -> np.ndarray:
    """glm::lookAtLH as a row-major 4x4."""
    eye = np.asarray(eye, float)
    f = np.asarray(center, float) - eye
    f /= np.linalg.norm(f)
    s = np.cross(up, f)
    s /= np.linalg.norm(s)
    u = np.cross(f, s)
    m = np.eye(4)
    m[0, :3] = s
    m[1, :3] = u
    m[2, :3] = f
    m[0, 3] = -s @ eye
    m[1, 3] = -u @ eye
    m[2, 3] = -f @ eye
    return m


@dataclass
class Camera:
    eye: np.ndarray          # float32 [3]
    inv_pv: np.ndarray       # float32 [16], column-major
    pv: np.ndarray           # float64 [4,4] row-major (for tests)
    width: int
    height: int
    fovy: float
    near: float
    far: float
    view: np.ndarray = None  # float32 [16], column-major View (the depth pass's uniform)
    proj: np.ndarray = None  # float32 [16], column-major Projection

    @property
    def inv_pv_colmajor(self) -> np.ndarray:
        return self.inv_pv


def make_camera(eye, target, width: int, height: int, fovy_deg: float = 60.0, near: float = 0.1,
                far: float = 1000.0, up=(0.0, 1.0, 0.0)) -> Camera:
    """Projection = perspective * scale(1,-1,1) (Camera3D.cpp:11), View = lookAtLH."""
    P = perspective_lh_zo(math.radians(fovy_deg), width / height, near, far)
    P = P @ np.diag([1.0, -1.0, 1.0, 1.0])
    V = look_at_lh(eye, target, np.asarray(up, float))
    PV = P @ V
    inv = np.linalg.inv(PV)
    inv_cm = np.ascontiguousarray(inv.T.reshape(16), dtype=np.float32)  # column-major
    cm = lambda m: np.ascontiguousarray(m.T.reshape(16), dtype=np.float32)
    return Camera(np.asarray(eye, np.float32), inv_cm, PV, width, height, fovy_deg, near, far, cm(V), cm(P))


def orbit_eye(target, radius: float, elevation_deg: float, azimuth_deg: float) -> np.ndarray:
    """Eye on a sphere around target: azimuth measured from -z towards -x, y up."""
    el, az = math.radians(elevation_deg), math.radians(azimuth_deg)
    t = np.asarray(target, float)
    return t + radius * np.array([-math.cos(el) * math.sin(az), math.sin(el), -math.cos(el) * math.cos(az)])


def project_depth(cam: Camera, pts: np.ndarray) -> np.ndarray:
    """NDC depth of world points (for reference-mode depth inputs in tests)."""
    ph = np.c_[pts, np.ones(len(pts))] @ cam.pv.T
    return ph[:, 2] / ph[:, 3]
