"""The 24 inputs of the step-proposal network at a march sample, restated
with numpy from the particle positions alone (test helper: the GPU trace rows
are checked against it)."""
import math

import numpy as np


def _keys(x, h):
    """Cell keys, the search grid's rule: int(x * (1/h)) in FP32, minus one for
    negative x."""
    v = (np.float32(1.0) / np.float32(h)) * np.asarray(x, np.float32)
    k = np.trunc(v.astype(np.float32)).astype(np.int64)
    return np.where(np.asarray(x) < 0, k - 1, k)


def point(a, s, d):
    """a + s d in FP32 with the product rounded first (no fused multiply-add):
    the sample positions of the march, whose kernels are built without
    contraction."""
    return (a + (np.float32(s) * d).astype(np.float32)).astype(np.float32)


class Features:
    """The 24 network inputs at a march sample restated with numpy from the
    particle positions alone: cell counts by brute-force keys, density and
    gradient in FP64 over every particle (the cubic spline of Kernel.cpp:16-32
    and its derivative), direction, slab distances and the 27-cell count."""

    def __init__(self, xyz, h, iso, origin, dims):
        self.xyz, self.h, self.iso = xyz.astype(np.float64), float(h), float(iso)
        self.origin, self.dims = np.asarray(origin, np.int64), np.asarray(dims, np.int64)
        self.k = _keys(xyz, h)
        self.fine = {}
        for kk in map(tuple, self.k):
            self.fine[kk] = self.fine.get(kk, 0) + 1
        self.coarse = {}
        for kk in map(tuple, (self.k - self.origin) >> 2):
            self.coarse[kk] = self.coarse.get(kk, 0) + 1

    def cell(self, p):
        return self.fine.get(tuple(_keys(p, self.h)), 0)

    def coarse_cell(self, p):
        c = _keys(p, self.h) - self.origin
        if np.any(c < 0) or np.any(c >= self.dims):
            return 0
        return self.coarse.get(tuple(c >> 2), 0)

    def count27(self, p):
        return int(np.all(np.abs(self.k - _keys(p, self.h)) <= 1, axis=1).sum())

    def rho_grad(self, q):
        r = self.xyz - np.asarray(q, np.float64)        # x_j - q
        dist = np.sqrt((r * r).sum(1))
        m = dist < self.h
        qq, rr, dd = dist[m] / self.h, r[m], dist[m]
        sig = 8.0 / (math.pi * self.h ** 3)
        w = np.where(qq <= 0.5, 6.0 * qq ** 3 - 6.0 * qq ** 2 + 1.0, 2.0 * (1.0 - qq) ** 3) * sig
        dw = np.where(qq <= 0.5, 18.0 * qq ** 2 - 12.0 * qq, -6.0 * (1.0 - qq) ** 2) * sig / self.h
        nz = dd > 0
        # d/dq W(|x_j - q|) = -W'(r) (x_j - q) / r
        g = (-(dw[nz] / dd[nz])[:, None] * rr[nz]).sum(0)
        scale = (np.abs(dw[nz])[:, None] * np.abs(rr[nz]) / dd[nz][:, None]).sum(0)
        return w.sum(), g, scale

    def row(self, o, d, t, tn, tf):
        """Expected features at sample t of the ray (o, d) entering the grid's
        box at tn and leaving it at tf."""
        f = np.zeros(24)
        p = point(o, t, d)
        for j in range(8):
            f[j] = min(self.cell(point(p, np.float32(j + 1) * np.float32(self.h), d)), 32) / 8.0
        for j in range(8):
            s = np.float32(j) * np.float32(4.0) * np.float32(self.h)
            f[8 + j] = min(self.coarse_cell(point(p, s, d)), 1024) / 256.0
        f[16:19] = d
        f[19] = min((tf - t) / self.h, 256.0) / 64.0
        rho, g, _ = self.rho_grad(p)
        f[20] = min(rho / self.iso, 4.0)
        f[21] = min(np.linalg.norm(g) * self.h / self.iso, 16.0) / 4.0
        f[22] = min((t - tn) / self.h, 256.0) / 64.0
        f[23] = min(self.count27(p), 512) / 128.0
        return f
