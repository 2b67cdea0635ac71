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
        # atan2(|a x b|, a.b): well conditioned for small angles, unlike arccos of the dot
        ang = np.degrees(np.arctan2(np.linalg.norm(np.cross(na, nb), axis=-1), (na * nb).sum(-1)))
        out.update(depth_max=float(dt.max()), depth_mean=float(dt.mean()), depth_p999=float(np.quantile(dt, 0.999)),
                   normal_max_deg=float(ang.max()), normal_p999_deg=float(np.quantile(ang, 0.999)))
    else:
        out.update(depth_max=0.0, depth_mean=0.0, depth_p999=0.0, normal_max_deg=0.0, normal_p999_deg=0.0)
    return out


def assert_within(m, ext, normal_key="normal_max_deg"):
    """Every pixel both marches hit is within the depth and normal tolerances."""
    assert m["mask_agree"] >= MASK_MIN, m
    assert m["depth_max"] <= DEPTH_REL * ext, (m, DEPTH_REL * ext)
    assert m[normal_key] <= NORMAL_DEG, m


def depth_in_front(cam, bmin, bmax, margin):
    """A depth buffer for the reference march: the NDC depth of the point `margin`
    in front of each ray's entry into [bmin, bmax] (1.0, the clear value, where the
    ray misses the box).  Rays follow the pixel-corner convention."""
    W, H = cam.width, cam.height
    M = cam.inv_pv.reshape(4, 4).T.astype(np.float64)
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    cx, cy = xs * 2.0 / W - 1.0, ys * 2.0 / H - 1.0
    ph = np.einsum("ij,jhw->ihw", M, np.stack([cx, cy, np.ones_like(cx), np.ones_like(cx)]))
    p = ph[:3] / ph[3]
    o = cam.eye.astype(np.float64)[:, None, None]
    d = p - o
    d /= np.linalg.norm(d, axis=0)
    lo = np.asarray(bmin, np.float64)[:, None, None]
    hi = np.asarray(bmax, np.float64)[:, None, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        t0, t1 = (lo - o) / d, (hi - o) / d
    tn = np.maximum(np.nanmax(np.minimum(t0, t1), axis=0), 0.0)
    tf = np.nanmin(np.maximum(t0, t1), axis=0)
    q = o + (tn - margin) * d
    qh = np.einsum("ij,jhw->ihw", cam.pv, np.concatenate([q, np.ones((1, H, W))]))
    z = (qh[2] / qh[3]).astype(np.float32)
    return np.where(tn <= tf, z, np.float32(1.0)).astype(np.float32)
