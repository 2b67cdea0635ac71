"""Seeded synthetic particle scenes and cameras for the march workloads.

The reference loads Houdini .bgeo frames (src/app/Dataset.cpp:169-227, 292-306)
and has no deterministic scene generator (its only RNG is the unseeded
RandomFloat of src/app/Utils.cpp:38-45), so the workloads are defined here.
Every scene is a float32 array of shape [N, 3]; h is the SPH support, which the
reference calls ParticleRadius (assets/config.yml:19, RayMarcher.cpp:98).

Kernel/iso conventions shared by all scenes: support h = 4 r, lattice spacing
2 r = h / 2, rest density rho0 = sigma * 3.1415061532636885 (the perfect-lattice
sum), iso = 0.5 * rho0, march step r / 2 = h / 8, central-difference offset
r / 8 = h / 32, 8 bisection iterations.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .camera import Camera, make_camera, orbit_eye

LATTICE_SUM = 3.1415061532636885


def sigma(h: float) -> float:
    """8 / (pi h^3) in float32, evaluated as the kernel does (Kernel.cpp:13)."""
    h = np.float32(h)
    return float(np.float32(8.0) / (np.float32(math.pi) * h * h * h))


def iso_for(h: float) -> float:
    return float(np.float32(0.5 * sigma(h) * LATTICE_SUM))


@dataclass
class Scene:
    name: str
    xyz: np.ndarray
    h: float
    iso: float
    step: float
    normal_eps: float
    cameras: list = field(default_factory=list)
    notes: str = ""

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    def settings(self, **kw) -> dict:
        d = dict(mode=0, max_steps=0, step=self.step, iso=self.iso, bisect_iters=8, final_fp64=1,
                 normal_eps=self.normal_eps, normal_mode=0, skip_mode=0, compensated=1)
        d.update(kw)
        return d


def _defaults(h: float):
    h = float(np.float32(h))
    return dict(h=h, iso=iso_for(h), step=float(np.float32(h / 8.0)), normal_eps=float(np.float32(h / 32.0)))


def lattice_cube(seed: int = 0, n_side: int = 32, spacing: float = 0.05, centre=(0.5, 0.5, 0.5),
                 jitter: float = 0.1, width: int = 640, height: int = 480) -> Scene:
    """32^3 lattice of spacing 2r = 0.05 centred on (0.5,0.5,0.5), uniform jitter of
    +-0.1 spacing, h = 4r = 0.1.  The 31*0.05 = 1.55 span overhangs the unit box.
    Cameras: 60 deg, distance 2 from the centre, one oblique and one axis view."""
    rng = np.random.default_rng(seed)
    i = np.arange(n_side, dtype=np.float64) - (n_side - 1) / 2.0
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    c = np.asarray(centre, float)
    xyz = c + g * spacing + rng.uniform(-jitter * spacing, jitter * spacing, g.shape)
    prm = _defaults(2.0 * spacing)
    d1 = np.array([0.45, 0.35, -0.82])
    d1 /= np.linalg.norm(d1)
    cams = [make_camera(c + 2.0 * d1, c, width, height),
            make_camera(c + np.array([0.0, 0.0, -2.0]), c, width, height)]
    return Scene("lattice_cube", xyz.astype(np.float32), cameras=cams,
                 notes=f"32^3 lattice, spacing {spacing}, jitter U(+-{jitter}*spacing), seed {seed}", **prm)


TANK = (1.0, 1.0, 0.4)
TANK_CENTRE = (0.5, 0.4, 0.2)
COLUMN_CENTRE = (0.2, 0.4, 0.2)


def dam_break(seed: int = 0, nx: int = 80, ny: int = 160, nz: int = 80, spacing: float = 0.005,
              jitter_sigma: float = 0.002, width: int = 1920, height: int = 1080,
              radius: float = 1.3, elevation: float = 30.0, azimuth: float = 45.0) -> Scene:
    """Dam-break column filling x in [0,0.4], y in [0,0.8], z in [0,0.4] of a 1x1x0.4 tank:
    80x160x80 = 1,024,000 particles at spacing 5 mm (r = 2.5 mm, h = 1 cm) with a
    seeded Gaussian jitter of sigma 2 mm.  Orbit camera at 30 deg elevation, 45 deg
    azimuth, radius 1.3, looking at the centre of the column (0.2, 0.4, 0.2)."""
    rng = np.random.default_rng(seed)
    ix = (np.arange(nx) + 0.5) * spacing
    iy = (np.arange(ny) + 0.5) * spacing
    iz = (np.arange(nz) + 0.5) * spacing
    g = np.stack(np.meshgrid(ix, iy, iz, indexing="ij"), -1).reshape(-1, 3)
    xyz = g + rng.normal(0.0, jitter_sigma, g.shape)
    prm = _defaults(2.0 * spacing)
    eye = orbit_eye(COLUMN_CENTRE, radius, elevation, azimuth)
    cams = [make_camera(eye, COLUMN_CENTRE, width, height)]
    return Scene("dam_break", xyz.astype(np.float32), cameras=cams,
                 notes=f"{nx}x{ny}x{nz} lattice spacing {spacing}, N(0,{jitter_sigma}) jitter, seed {seed}", **prm)


ANISO_SETTINGS = dict(anisotropic=1, k_n=0.5, k_r=2.0, k_s=2000.0, n_eps=1, ext_mult=2.0)
"""The reference's defaults (AdvancedRenderer.cpp:18-28, config.yml particleRadiusMultiplier)."""


def aniso_settings(sc: "Scene", iso: float = 1.0, **kw) -> dict:
    """Configured-march settings of the anisotropic density for a scene at the
    reference's tuning (h = 0.1, IsoDensity 1.0): the scene's step and normal
    offset, the reference's k_n, k_r, k_s, N_eps and radius multiplier."""
    d = sc.settings(iso=float(np.float32(iso)), **ANISO_SETTINGS)
    d.update(kw)
    return d


def dam_break_aniso(seed: int = 0, scale: float = 10.0, width: int = 1920, height: int = 1080, **kw) -> Scene:
    """The dam-break column scaled by 10 about the origin: 1,024,000 particles at
    spacing 5 cm, h = 0.1 -- the particle radius the reference's anisotropic
    march is tuned for (config.yml particleRadius 0.1; AdvancedRenderer.cpp:18-28
    k_s = 2000 against covariances of order h^2: G ~ I/h in the bulk, which
    makes the kernel scale-dependent).  Same camera geometry, scaled."""
    base = dam_break(seed, width=width, height=height, **kw)
    xyz = (base.xyz.astype(np.float64) * scale).astype(np.float32)
    prm = _defaults(float(np.float32(base.h * scale)))
    eye = orbit_eye(tuple(scale * c for c in COLUMN_CENTRE), 1.3 * scale, 30.0, 45.0)
    cams = [make_camera(eye, tuple(scale * c for c in COLUMN_CENTRE), width, height)]
    return Scene("dam_break_aniso", xyz, cameras=cams,
                 notes=f"dam_break(seed {seed}) scaled by {scale}: h = {prm['h']}", **prm)


def displaced_frames(base: Scene, frames: int, seed: int, rel: float = 0.01, spacing: float = 0.005):
    """Frame f = frame f-1 + an independent uniform displacement of norm <= rel*spacing
    per particle (seeded), i.e. a slowly moving particle set."""
    rng = np.random.default_rng(seed)
    out = [base.xyz]
    cur = base.xyz.astype(np.float64)
    for _ in range(frames - 1):
        v = rng.normal(size=cur.shape)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        mag = rel * spacing * rng.uniform(0.0, 1.0, (cur.shape[0], 1)) ** (1.0 / 3.0)
        cur = cur + v * mag
        out.append(cur.astype(np.float32))
    return out


def clustered_pool(seed: int = 0, n_total: int = 4_000_000, r: float = 0.003, width: int = 3840,
                   height: int = 2160, n_drops: int = 200) -> Scene:
    """Uneven-density scene in the 1x1x0.4 tank, h = 4r = 0.012.
    60% of the particles form a jittered pool layer y in [0, 0.15] (lattice spacing
    chosen so the layer holds that many), 40% are split evenly: half fill 200
    spherical droplets (radius U[0.01,0.05], centre uniform over the upper half),
    half lie in splash cones (apex at each droplet, axis down, 30 deg half-angle,
    length 3x the droplet radius)."""
    rng = np.random.default_rng(seed)
    n_pool = int(0.6 * n_total)
    vol = TANK[0] * 0.15 * TANK[2]
    s = (vol / n_pool) ** (1.0 / 3.0)
    nx, ny, nz = int(round(TANK[0] / s)), int(round(0.15 / s)), int(round(TANK[2] / s))
    while nx * ny * nz < n_pool:
        nx += 1
    ix = (np.arange(nx) + 0.5) * (TANK[0] / nx)
    iy = (np.arange(ny) + 0.5) * (0.15 / ny)
    iz = (np.arange(nz) + 0.5) * (TANK[2] / nz)
    g = np.stack(np.meshgrid(ix, iy, iz, indexing="ij"), -1).reshape(-1, 3)
    g = g[rng.permutation(g.shape[0])[:n_pool]]
    pool = g + rng.uniform(-0.1 * s, 0.1 * s, g.shape)
    n_rest = n_total - n_pool
    n_drop = n_rest // 2
    n_cone = n_rest - n_drop
    radii = rng.uniform(0.01, 0.05, n_drops)
    centres = np.c_[rng.uniform(0.05, 0.95, n_drops), rng.uniform(0.5, 0.95, n_drops), rng.uniform(0.05, 0.35, n_drops)]
    w = radii ** 3 / np.sum(radii ** 3)
    counts = rng.multinomial(n_drop, w)
    drops = []
    for c, rr, k in zip(centres, radii, counts):
        v = rng.normal(size=(k, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        drops.append(c + v * rr * rng.uniform(0, 1, (k, 1)) ** (1.0 / 3.0))
    ccount = rng.multinomial(n_cone, np.full(n_drops, 1.0 / n_drops))
    cones = []
    for c, rr, k in zip(centres, radii, ccount):
        L = 3.0 * rr
        t = L * rng.uniform(0, 1, k) ** (1.0 / 3.0)
        ang = rng.uniform(0, 2 * math.pi, k)
        rad = t * math.tan(math.radians(30.0)) * np.sqrt(rng.uniform(0, 1, k))
        cones.append(np.c_[c[0] + rad * np.cos(ang), c[1] - rr - t, c[2] + rad * np.sin(ang)])
    xyz = np.concatenate([pool] + drops + cones).astype(np.float32)
    prm = _defaults(4.0 * r)
    eye = orbit_eye((0.5, 0.3, 0.2), 1.7, 30.0, 45.0)
    cams = [make_camera(eye, (0.5, 0.3, 0.2), width, height)]
    return Scene("clustered_pool", xyz, cameras=cams,
                 notes=f"{n_total} particles: pool {n_pool} at spacing {s:.5f}, {n_drops} droplets, cones; seed {seed}",
                 **prm)


def tiny_scene(points, h: float, eye, target, width=64, height=48) -> Scene:
    prm = _defaults(h)
    cams = [make_camera(np.asarray(eye, float), np.asarray(target, float), width, height)]
    return Scene("tiny", np.asarray(points, np.float32).reshape(-1, 3), cameras=cams, **prm)


def random_views(rng: np.random.Generator, target, n: int, wmin=256, wmax=1920, hmin=256, hmax=1080,
                 radius=(1.3, 2.0)) -> list[Camera]:
    """n cameras with W ~ U{wmin..wmax}, H ~ U{hmin..hmax}, orbiting target."""
    cams = []
    for _ in range(n):
        W = int(rng.integers(wmin, wmax + 1))
        H = int(rng.integers(hmin, hmax + 1))
        eye = orbit_eye(target, float(rng.uniform(*radius)), float(rng.uniform(10, 60)), float(rng.uniform(0, 360)))
        cams.append(make_camera(eye, target, W, H))
    return cams
