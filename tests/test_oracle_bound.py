"""The sub-box density bound of the configured march's skip, on the CPU oracle.

Pins: (1) safety -- points anywhere in a flagged sub-box, faces and corners
included, have oracle density below iso; (2) the bits against a float64
brute force written independently here (sub-box geometry, stencil and
sub-index layout l = sx + 4 sy + 16 sz); (3) the march with the skip gives the
image of the march without it, bit for bit, with fewer march samples."""
import numpy as np
import pytest

import oracle as O
from pmarch import scenes

LIM = 1.0 - 2.0 ** -9


def _sub_boxes(G, bits):
    """(cell dense coords, sub index) of every set bit in non-trivial cells."""
    full = np.uint64(0xFFFFFFFFFFFFFFFF)
    nx, ny, nz = G.dims
    out = []
    for c in np.nonzero((bits != full) & (bits != 0))[0]:
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        b = int(bits[c])
        out += [((x, y, z), l) for l in range(64) if (b >> l) & 1]
    return out


def _points_in(G, cell, l, rng, n):
    """Points inside sub-box l of a dense cell (u = x / h in [key + s/4, key + (s+1)/4)),
    corners and face centres included."""
    h = np.float32(G.h)
    key = np.array(cell) + np.array(G.origin)
    s = np.array([l & 3, (l >> 2) & 3, l >> 4])
    lo = (key + s * 0.25) * np.float64(h)
    hi = (key + (s + 1) * 0.25) * np.float64(h)
    corners = np.array([[lo[0] if i & 1 else hi[0], lo[1] if i & 2 else hi[1], lo[2] if i & 4 else hi[2]]
                        for i in range(8)])
    inner = lo + (hi - lo) * rng.uniform(0, 1, (n, 3))
    pts = np.concatenate([corners * (1 - 1e-7), inner]).astype(np.float32)
    return pts


@pytest.mark.parametrize("scene", ["lattice", "clustered"])
def test_bound_is_safe(scene):
    if scene == "lattice":
        sc = scenes.lattice_cube(3, n_side=14)
    else:
        sc = scenes.clustered_pool(5, n_total=6000, r=0.02, width=64, height=48)
    G = O.Grid(sc.xyz, sc.h)
    bits = G.bound(sc.iso)
    boxes = _sub_boxes(G, bits)
    assert len(boxes) > 500
    rng = np.random.default_rng(0)
    pick = rng.choice(len(boxes), size=min(1500, len(boxes)), replace=False)
    pts = np.concatenate([_points_in(G, *boxes[i], rng, 6) for i in pick])
    rho = G.density(pts, compensated=True)
    print(scene, "flagged sub-boxes", len(boxes), "points", len(pts), "max rho / iso", rho.max() / sc.iso)
    assert rho.max() < sc.iso


def _brute_bits(G, iso, xyz):
    """float64 restatement: U = sum_j W(dist(x_j, widened sub-box)) over particles
    within h of the box, bit set when U < iso (1 - 2^-9)."""
    h = float(G.h)
    sig = 8.0 / (np.pi * h ** 3)
    nx, ny, nz = G.dims
    ox, oy, oz = G.origin
    es = G.empty27()
    bits = np.zeros(nx * ny * nz, np.uint64)
    U = np.zeros((nx * ny * nz, 64))
    sx = np.arange(64) & 3
    sy = (np.arange(64) >> 2) & 3
    sz = np.arange(64) >> 4
    p = xyz.astype(np.float64)
    for c in range(nx * ny * nz):
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        if min(x, y, z) < 1 or x > nx - 2 or y > ny - 2 or z > nz - 2 or es[c]:
            bits[c] = np.uint64(0xFFFFFFFFFFFFFFFF)
            U[c] = -1
            continue
        key = np.array([x + ox, y + oy, z + oz], np.float64)
        km = np.abs(key).max() + 1
        delta = h * (2.0 ** -10 + km * 2.0 ** -18)
        s = np.stack([sx, sy, sz], 1)
        lo = (key + 0.25 * s) * h - delta
        hi = (key + 0.25 * (s + 1)) * h + delta
        e = np.maximum(np.maximum(lo[:, None, :] - p[None], p[None] - hi[:, None, :]), 0.0)
        r = np.sqrt((e ** 2).sum(-1)) / h
        w = np.where(r < 0.5, 6.0 * (r ** 3 - r ** 2) + 1.0, 2.0 * np.clip(1.0 - r, 0, None) ** 3) * sig
        U[c] = w.sum(1)
        b = 0
        for l in range(64):
            if U[c, l] < iso * LIM:
                b |= 1 << l
        bits[c] = np.uint64(b)
    return bits, U


def test_bound_bits_match_fp64_bruteforce():
    rng = np.random.default_rng(11)
    # a small irregular blob: dense core, sparse halo
    core = rng.normal(0.0, 0.05, (500, 3))
    halo = rng.uniform(-0.2, 0.2, (150, 3))
    xyz = np.concatenate([core, halo]).astype(np.float32)
    h = 0.05
    iso = scenes.iso_for(h)
    G = O.Grid(xyz, h)
    bits = G.bound(iso)
    ref, U = _brute_bits(G, iso, xyz)
    diff = np.nonzero(bits != ref)[0]
    # only sub-boxes whose bound sits within float rounding of the limit may differ
    for c in diff:
        d = int(bits[c]) ^ int(ref[c])
        for l in range(64):
            if (d >> l) & 1:
                assert abs(U[c, l] - iso * LIM) < 1e-5 * iso, (c, l, U[c, l])
    nontrivial = np.sum((ref != np.uint64(0xFFFFFFFFFFFFFFFF)) & (ref != 0))
    print("cells", bits.size, "non-trivial", nontrivial, "differing", diff.size)
    assert nontrivial > 50


@pytest.mark.parametrize("scene", ["lattice", "clustered"])
def test_bound_skip_march_identical(scene):
    if scene == "lattice":
        sc = scenes.lattice_cube(1, width=128, height=96)
    else:
        sc = scenes.clustered_pool(2, n_total=20000, r=0.01, width=128, height=96)
    G = O.Grid(sc.xyz, sc.h)
    c = sc.cameras[0]
    keys = dict(O.Settings._fields_)
    out = {}
    for sk in (0, 4):
        st = O.make_settings(**{k: v for k, v in sc.settings(skip_mode=sk).items() if k in keys})
        out[sk] = O.march(G, st, c.inv_pv, c.eye, c.width, c.height)
    for a, b in zip(out[0][:3], out[4][:3]):
        assert np.array_equal(a, b)
    c0, c4 = out[0][3], out[4][3]
    print(scene, "march samples", c0["march_samples"], "->", c4["march_samples"])
    assert c4["hits"] == c0["hits"] and c4["hits"] > 0
    assert c4["march_samples"] < c0["march_samples"]
    assert c4["samples"] - c4["march_samples"] == c0["samples"] - c0["march_samples"]


def test_bound_march_needs_build():
    sc = scenes.lattice_cube(0, n_side=8, width=16, height=12)
    G = O.Grid(sc.xyz, sc.h)
    keys = dict(O.Settings._fields_)
    st = O.make_settings(**{k: v for k, v in sc.settings(skip_mode=4).items() if k in keys})
    c = sc.cameras[0]
    O.march(G, st, c.inv_pv, c.eye, 16, 12)      # the wrapper builds the bits
    st.iso = st.iso * 2
    with pytest.raises(RuntimeError):
        pos = np.zeros((12, 16, 4), np.float32)
        rc = O.lib().or_march(O.C.byref(G.g), O.C.byref(st), np.ascontiguousarray(c.inv_pv, np.float32).reshape(16),
                              np.ascontiguousarray(c.eye, np.float32), 16, 12, None, 1, 1, pos.reshape(-1),
                              pos.copy().reshape(-1), None, O.C.byref(O.Counters()))
        if rc != 0:
            raise RuntimeError(rc)
