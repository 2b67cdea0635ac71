"""Pins of the fused central-difference walk's cells (or_cdiff_union): the walk
around q with the box [q - eps, q + eps] must visit, in the march grid's
order, every particle within h of each of the six samples q +- eps e_a (so a
sum over its list filtered per sample is that sample's single-sample sum), and
list nothing beyond (h + eps)(1 + 2^-10) of q -- checked against brute force
in float32 and against the single-sample neighbour lists.  No GPU."""
import numpy as np
import pytest

import oracle as O
from pmarch import scenes


def _six(q, eps):
    q = np.asarray(q, np.float32)
    e = np.float32(eps)
    pts = []
    for a in range(3):
        for sgn in (1, -1):
            p = q.copy()
            p[a] = np.float32(q[a] + e) if sgn > 0 else np.float32(q[a] - e)
            pts.append(p)
    return pts


def _r2(xyz, q):
    r = xyz.astype(np.float32) - np.asarray(q, np.float32)
    return (r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2]


def _is_subsequence(sub, seq):
    it = iter(map(tuple, seq))
    return all(any(t == s for s in it) for t in map(tuple, sub))


def _queries(G, xyz, h, rng, n):
    """Points near particles, half of them moved to within 1/20 fine cell of a
    fine-cell boundary on a random axis (the cases where the box spans two
    fine cells)."""
    idx = rng.integers(0, len(xyz), n)
    q = xyz[idx] + rng.uniform(-h, h, (n, 3)).astype(np.float32)
    hf = h / 2.0
    for i in range(0, n, 2):
        a = rng.integers(0, 3)
        k = np.floor(q[i, a] / hf)
        q[i, a] = np.float32((k + 1.0 + rng.uniform(-0.05, 0.05)) * hf)
    return q.astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1])
def test_union_covers_six_samples(seed):
    sc = scenes.dam_break(seed, nx=16, ny=24, nz=16)
    xyz, h = sc.xyz, sc.h
    eps = sc.normal_eps
    G = O.Grid(xyz, h)
    rng = np.random.default_rng(seed + 10)
    h2 = np.float32(h) * np.float32(h)
    he = np.float32(np.float32(np.float32(h) + np.float32(eps)) * np.float32(1.0009765625))
    R2 = he * he
    spans = 0
    for q in _queries(G, xyz, h, rng, 300):
        res = G.cdiff_union(q, eps)
        assert res is not None
        cand, lst = res
        assert cand >= len(lst)
        # nothing beyond (h + eps)(1 + 2^-10) of q
        if len(lst):
            assert np.all(_r2(lst, q) < R2)
        # every particle within h of a sample is listed, in the sample's own order
        for p in _six(q, eps):
            nb = xyz[_r2(xyz, p) < h2]
            got = lst[_r2(lst, p) < h2] if len(lst) else lst
            assert len(got) == len(nb)
            own, _ = G.neighbors_fine(p)
            assert np.array_equal(own, got)
            assert _is_subsequence(own, lst)
        f = np.float32(2.0 / h) * q
        spans += int(np.any(np.floor(f - np.float32(2.0 / h) * np.float32(eps)) != np.floor(f + np.float32(2.0 / h) * np.float32(eps))))
    assert spans > 100          # the box crossed a fine-cell boundary in many queries


def test_union_leaves_grid_far_outside():
    sc = scenes.lattice_cube(0)
    G = O.Grid(sc.xyz, sc.h)
    far = np.asarray(G.bmax, np.float32) + np.float32(40 * sc.h)
    assert G.cdiff_union(far, sc.normal_eps) is None
