"""The sub-box density bound of the configured march's skip, on the CPU oracle.

Pins: (1) safety -- points anywhere in a flagged sub-box, faces and corners
included, have oracle density below iso; (2) the bits against a float64
brute force written independently here (sub-box geometry 8x8x8 per cell, the
particles near the box, and the bit layout: sub-box (fx, fy, fz) is bit
8 l + (fx & 1) + 2 (fy & 1) + 4 (fz & 1) with l = fx/2 + 4 (fy/2) + 16 (fz/2));
(3) the march with the skip gives the image of the march without it, bit for
bit, with fewer march samples."""
import numpy as np
import pytest

import oracle as O
from pmarch import scenes

LIM = 1.0 - 2.0 ** -9
FULL = np.uint64(0xFFFFFFFFFFFFFFFF)


def _fine_index(b):
    """Per-axis eighth-box indices of bit b of a cell's 512."""
    l, f = b >> 3, b & 7
    return np.array([2 * (l & 3) + (f & 1), 2 * ((l >> 2) & 3) + ((f >> 1) & 1), 2 * (l >> 4) + (f >> 2)])


def _sub_boxes(G, bits):
    """(cell dense coords, bit) of every set bit in non-trivial cells."""
    nx, ny, nz = G.dims
    out = []
    trivial = np.all(bits == FULL, axis=1) | np.all(bits == 0, axis=1)
    for c in np.nonzero(~trivial)[0]:
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        out += [((x, y, z), b) for b in range(512) if (int(bits[c, b >> 6]) >> (b & 63)) & 1]
    return out


def _points_in(G, cell, b, rng, n):
    """Points inside sub-box b of a dense cell (u = x / h in [key + f/8, key + (f+1)/8)),
    corners included."""
    h = np.float32(G.h)
    key = np.array(cell) + np.array(G.origin)
    f = _fine_index(b)
    lo = (key + f * 0.125) * np.float64(h)
    hi = (key + (f + 1) * 0.125) * np.float64(h)
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
    """float64 restatement: U = sum_j W(dist(x_j, widened box)) over the
    particles within h of the cell, for the 64 quarter boxes (4x4x4) and the
    512 sub-boxes (8x8x8); a sub-box's bit is set when its quarter box's U, or
    else its own, is below the limit iso (1 - 2^-9), except in cells where the
    particles of the centre row (keys x-1..x+1 at the cell's y and z) alone
    take every quarter box's U to the limit, and in quarter boxes whose U
    reaches 4 x the limit (those get no bit).  Returns the bits, per sub-box
    its own U and its quarter box's U, and per cell the least centre-row U."""
    h = float(G.h)
    sig = 8.0 / (np.pi * h ** 3)
    nx, ny, nz = G.dims
    ox, oy, oz = G.origin
    es = G.empty27()
    bits = np.zeros((nx * ny * nz, 8), np.uint64)
    U = np.zeros((nx * ny * nz, 512))
    UQ = np.zeros((nx * ny * nz, 512))
    f = np.stack([_fine_index(b) for b in range(512)])
    p_all = xyz.astype(np.float64)
    inv = np.float32(1.0) / np.float32(h)
    v = (inv * xyz.astype(np.float32)).astype(np.float32)
    keys = np.where(xyz < 0, np.trunc(v).astype(np.int64) - 1, np.trunc(v).astype(np.int64))
    ucr = np.full(nx * ny * nz, np.inf)             # least centre-row U of a quarter box
    for c in range(nx * ny * nz):
        x, y, z = c % nx, (c // nx) % ny, c // (nx * ny)
        if min(x, y, z) < 1 or x > nx - 2 or y > ny - 2 or z > nz - 2 or es[c]:
            bits[c] = FULL
            U[c] = -1
            continue
        key = np.array([x + ox, y + oy, z + oz], np.float64)
        near = np.all((p_all > key * h - 1.01 * h) & (p_all < (key + 1) * h + 1.01 * h), axis=1)
        km = np.abs(key).max() + 1
        delta = h * (2.0 ** -10 + km * 2.0 ** -18)

        def box_u(p, lo, hi):
            e = np.maximum(np.maximum(lo[:, None, :] - p[None], p[None] - hi[:, None, :]), 0.0)
            r = np.sqrt((e ** 2).sum(-1)) / h
            w = np.where(r < 0.5, 6.0 * (r ** 3 - r ** 2) + 1.0, 2.0 * np.clip(1.0 - r, 0, None) ** 3) * sig
            return w.sum(1)
        q = f // 2                                   # the quarter box of each sub-box
        qlo, qhi = (key + 0.25 * q) * h - delta, (key + 0.25 * (q + 1)) * h + delta
        kc = key.astype(np.int64)
        row = (np.abs(keys[:, 0] - kc[0]) <= 1) & (keys[:, 1] == kc[1]) & (keys[:, 2] == kc[2])
        ucr[c] = box_u(p_all[row], qlo, qhi).min()
        if ucr[c] >= iso * LIM:
            U[c] = np.inf                            # inside the fluid: no bit
            continue
        uf = box_u(p_all[near], (key + 0.125 * f) * h - delta, (key + 0.125 * (f + 1)) * h + delta)
        uq = box_u(p_all[near], qlo, qhi)
        U[c], UQ[c] = uf, uq
        words = [0] * 8
        for b in range(512):
            if uq[b] < iso * LIM or (uq[b] < 4 * iso * LIM and uf[b] < iso * LIM):
                words[b >> 6] |= 1 << (b & 63)
        bits[c] = np.array(words, np.uint64)
    return bits, U, UQ, ucr


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
    assert bits.shape == (int(np.prod(G.dims)), 8)
    ref, U, UQ, ucr = _brute_bits(G, iso, xyz)
    diff = np.nonzero(np.any(bits != ref, axis=1))[0]
    # only sub-boxes whose bounds (or whose cell's inside test) sit within
    # float rounding of the thresholds may differ
    lim, tol = iso * LIM, 1e-5 * iso
    for c in diff:
        if abs(ucr[c] - lim) < tol:
            continue
        for w in range(8):
            d = int(bits[c, w]) ^ int(ref[c, w])
            for k in range(64):
                if (d >> k) & 1:
                    b = 64 * w + k
                    near = (abs(UQ[c, b] - lim) < tol or abs(UQ[c, b] - 4 * lim) < 4 * tol
                            or abs(U[c, b] - lim) < tol)
                    assert near, (c, b, U[c, b], UQ[c, b])
    print("cells inside (no bit)", int(np.sum(np.isinf(U[:, 0]))),
          "quarter boxes at 4x the limit", int(np.sum(UQ[:, ::8] >= 4 * lim)))
    nontrivial = np.sum(~(np.all(ref == FULL, axis=1) | np.all(ref == 0, axis=1)))
    # the quarter boxes (bytes) that the march crosses in one jump
    quarters = sum(bin(int(v)).count("1") for v in bits.view(np.uint8).reshape(-1) if v == 0xFF)
    print("cells", len(bits), "non-trivial", nontrivial, "differing", diff.size, "full bytes", quarters // 8)
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


def test_quarter_boxes_level():
    """bound_boxes 1: every byte 0xff or 0; its set bits are set at the 8x8x8
    level too; the march's image is unchanged and its march samples lie between
    those of the 8x8x8 level and those of no skip."""
    sc = scenes.clustered_pool(2, n_total=20000, r=0.01, width=128, height=96)
    G = O.Grid(sc.xyz, sc.h)
    q = G.bound(sc.iso, 1)
    b = q.view(np.uint8)
    assert np.all((b == 0) | (b == 0xFF))
    f = G.bound(sc.iso, 0)
    assert np.array_equal(q & f, q)
    assert np.sum(f != q) > 0
    c = sc.cameras[0]
    keys = dict(O.Settings._fields_)
    out = {}
    for sk, boxes in ((0, 0), (4, 0), (4, 1)):
        st = O.make_settings(**{k: v for k, v in sc.settings(skip_mode=sk, bound_boxes=boxes).items() if k in keys})
        out[(sk, boxes)] = O.march(G, st, c.inv_pv, c.eye, c.width, c.height)
    for a, b4, b1 in zip(out[(0, 0)][:3], out[(4, 0)][:3], out[(4, 1)][:3]):
        assert np.array_equal(a, b4) and np.array_equal(a, b1)
    m = [out[k][3]["march_samples"] for k in ((0, 0), (4, 1), (4, 0))]
    print("march samples: no skip, 4x4x4, 8x8x8", m)
    assert m[0] > m[1] > m[2]


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
