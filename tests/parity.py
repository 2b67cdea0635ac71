"""Parity metrics between two marches (tolerances from the march's contract:
hit/miss mask agreement >= 99.9 %, depth error <= 1e-4 of the scene extent,
normal error <= 0.5 degrees)."""
import numpy as np

MASK_MIN = 0.999
DEPTH_REL = 1e-4
NORMAL_DEG = 0.5


def extent(xyz) -> float:
    """Longest side of the particle AABB."""
    xyz = np.asarray(xyz, np.float64)
    return float((xyz.max(0) - xyz.min(0)).max())


def compare(pos_a, nrm_a, t_a, pos_b, nrm_b, t_b):
    ma, mb = pos_a[..., 3] > 0, pos_b[..., 3] > 0
    both = ma & mb
    out = {"mask_agree": float((ma == mb).mean()), "hits_a": int(ma.sum()), "hits_b": int(mb.sum()),
           "both": int(both.sum())}
    if both.any():
        dt = np.abs(t_a[both].astype(np.float64) - t_b[both].astype(np.float64))
        na = nrm_a[both][:, :3].astype(np.float64)
        nb = nrm_b[both][:, :3].astype(np.float64)
        cosv = np.clip((na * nb).sum(-1) / np.maximum(np.linalg.norm(na, axis=-1) * np.linalg.norm(nb, axis=-1), 1e-300), -1, 1)
        ang = np.degrees(np.arccos(cosv))
        out.update(depth_max=float(dt.max()), depth_mean=float(dt.mean()), normal_max_deg=float(ang.max()),
                   normal_p999_deg=float(np.quantile(ang, 0.999)))
    else:
        out.update(depth_max=0.0, depth_mean=0.0, normal_max_deg=0.0, normal_p999_deg=0.0)
    return out


def assert_within(m, ext, normal_key="normal_max_deg"):
    assert m["mask_agree"] >= MASK_MIN, m
    assert m["depth_max"] <= DEPTH_REL * ext, (m, DEPTH_REL * ext)
    assert m[normal_key] <= NORMAL_DEG, m
