"""Pins of the oracle's uniform grid and neighbour search (Dataset.cpp:49-76, 272-280;
the CompactNSearch point query) against brute force and numpy -- no GPU."""
import numpy as np
import pytest

import oracle as O


def np_keys(xyz, h):
    """int(x / h), minus one for negative x, in float32 (x * (1/h))."""
    inv = np.float32(1.0) / np.float32(h)
    v = (inv * xyz.astype(np.float32)).astype(np.float32)
    k = np.trunc(v).astype(np.int64)
    return np.where(xyz < 0, k - 1, k)


def test_keys_negative_and_zero():
    h = 0.1
    assert O.key(0.0, h) == 0
    assert O.key(-0.0, h) == 0
    assert O.key(0.15, h) == 1
    assert O.key(-0.05, h) == -1
    assert O.key(-0.1, h) == -2      # int(-1.0000000149) - 1
    assert O.key(-0.25, h) == -3
    rng = np.random.default_rng(0)
    x = rng.uniform(-2, 2, 2000).astype(np.float32)
    assert all(O.key(float(v), h) == k for v, k in zip(x, np_keys(x, h)))


@pytest.mark.parametrize("seed", [0, 1])
def test_counting_sort_layout(seed):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-0.3, 0.45, (6000, 3)).astype(np.float32)
    h = 0.07
    G = O.Grid(xyz, h)
    k = np_keys(xyz, h)
    org = np.array(G.origin)
    dims = np.array(G.dims)
    assert np.array_equal(org, k.min(0) - 2) and np.array_equal(dims, k.max(0) - k.min(0) + 5)
    c = k - org
    lin = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
    counts = np.bincount(lin, minlength=int(np.prod(dims)))
    cs = G.cell_start()
    assert np.array_equal(np.diff(cs), counts)
    order = np.lexsort((np.arange(len(xyz)), lin))      # by cell, then input index
    assert np.array_equal(G.sorted_index(), order)
    assert np.array_equal(G.sorted_xyz()[:, :3], xyz[order])
    # padded AABB = bounds -+ h (Frame::ComputeAABB)
    assert np.array_equal(G.bmin, (xyz.min(0) - np.float32(h)).astype(np.float32))
    assert np.array_equal(G.bmax, (xyz.max(0) + np.float32(h)).astype(np.float32))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neighbors_grid_hash_brute(seed):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-0.2, 0.3, (3000, 3)).astype(np.float32)
    h = 0.06
    G = O.Grid(xyz, h)
    q = rng.uniform(-0.3, 0.4, (300, 3)).astype(np.float32)
    q[:20] = xyz[:20]                       # queries exactly on particles
    for p in q:
        nb_g, cand = G.neighbors(p)
        nb_b = O.neighbors_brute(xyz, h, p)
        nb_h = O.neighbors_hash(xyz, h, p)
        assert np.array_equal(np.sort(nb_g), nb_b)          # same set as brute force
        assert np.array_equal(nb_g, nb_h)                   # same sequence as the hashed table
        # candidates = particles in the 27 cells around key(p)
        kp = np_keys(p[None], h)[0]
        kk = np_keys(xyz, h)
        assert cand == int(np.all(np.abs(kk - kp) <= 1, axis=1).sum())


def test_neighbors_strict_support():
    h = 0.125
    xyz = np.array([[0, 0, 0], [0.125, 0, 0], [0.0625, 0, 0]], np.float32)
    G = O.Grid(xyz, h)
    nb, _ = G.neighbors(np.zeros(3, np.float32))
    assert sorted(nb.tolist()) == [0, 2]            # |r| = h is excluded


def test_densitygrid_conservative_flag_clears_only_empty_space():
    rng = np.random.default_rng(3)
    xyz = np.concatenate([rng.uniform(0, 0.3, (2000, 3)), rng.uniform(0.7, 0.8, (300, 3))]).astype(np.float32)
    h = 0.05
    G = O.Grid(xyz, h)
    dims, count_box, flag_box, flag_cons = G.densitygrid()
    ext = G.bmax - G.bmin
    assert dims == tuple(int(np.ceil(np.float32(e) / np.float32(h))) for e in ext)
    assert flag_box.sum() <= flag_cons.sum()
    # sample points inside cleared cells: density exactly zero
    cleared = np.flatnonzero(flag_cons == 0)
    pick = rng.choice(cleared, 300, replace=False)
    z, rem = np.divmod(pick, dims[0] * dims[1])
    y, x = np.divmod(rem, dims[0])
    cell = np.stack([x, y, z], -1)
    pts = G.bmin + (cell + rng.uniform(0, 1, cell.shape)) * h
    assert np.all(G.density64(pts) == 0.0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fine_stencil_layout_and_neighbours(seed):
    """The march grid: fine keys are the h-keys' rule with 2/h, stored
    fine-cell-major by input index; the trimmed 5x5x5 fine stencil returns
    exactly the brute-force neighbour set, ordered by (fine cell, input index),
    and visits fewer candidates than the 27 h-cells."""
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-0.2, 0.3, (4000, 3)).astype(np.float32)
    xyz[:50] = np.round(xyz[:50] / 0.03) * 0.03               # points on fine-cell faces, negative ones included
    h = 0.06
    G = O.Grid(xyz, h)
    kf = np_keys(xyz, h / 2)
    inv_hf = np.float32(2.0) * (np.float32(1.0) / np.float32(h))
    assert np.float32(G.g.inv_hf) == inv_hf
    org = np.array(G.g.forigin[:])
    dims = np.array(G.g.fdims[:])
    assert np.array_equal(org, kf.min(0) - 4) and np.array_equal(dims, kf.max(0) - kf.min(0) + 9)
    c = kf - org
    lin = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
    order = np.lexsort((np.arange(len(xyz)), lin))
    fs = np.ctypeslib.as_array(G.g.fsorted, shape=(len(xyz), 4))
    assert np.array_equal(fs[:, :3], xyz[order])
    rank = np.empty(len(xyz), np.int64)
    rank[order] = np.arange(len(xyz))
    q = rng.uniform(-0.3, 0.4, (400, 3)).astype(np.float32)
    q[:30] = xyz[:30]
    q[30:60] = np.round(q[30:60] / 0.03) * 0.03
    tot_fine = tot_h = 0
    for p in q:
        pos, cand = G.neighbors_fine(p)
        nb = O.neighbors_brute(xyz, h, p)
        assert len(pos) == len(nb)
        want = xyz[nb[np.argsort(rank[nb])]]                 # brute-force set in (fine cell, index) order
        assert np.array_equal(pos, want)
        tot_fine += cand
        tot_h += G.neighbors(p)[1]
    print("candidates fine", tot_fine, "27 h-cells", tot_h)
    assert tot_fine < 0.6 * tot_h


def test_fine_density_matches_fp64_brute_force():
    rng = np.random.default_rng(4)
    xyz = rng.normal(0.0, 0.05, (6000, 3)).astype(np.float32)
    h = 0.04
    G = O.Grid(xyz, h)
    q = rng.normal(0.0, 0.06, (500, 3)).astype(np.float32)
    rho = G.density(q)
    d = np.linalg.norm(xyz[None, :, :].astype(np.float64) - q[:, None, :].astype(np.float64), axis=-1) / h
    sig = 8.0 / (np.pi * h ** 3)
    w = np.where(d < 0.5, 6.0 * (d ** 3 - d ** 2) + 1.0, 2.0 * np.clip(1.0 - d, 0.0, None) ** 3) * sig
    ref = w.sum(1)
    assert np.max(np.abs(rho - ref) / np.maximum(ref, sig)) < 2e-6


def test_per_sample_candidate_counts():
    """or_stencil_cand is the trimmed fine stencil's candidate count (the walk of
    or_neighbors_fine), and or_cand27 the number of particles whose search-cell
    key is within one of the point's on every axis (brute force), 0 where the
    3x3x3 block leaves the grid."""
    rng = np.random.default_rng(4)
    xyz = rng.uniform(-0.3, 0.5, (6000, 3)).astype(np.float32)
    h = np.float32(0.05)
    G = O.Grid(xyz, h)
    q = np.concatenate([rng.uniform(-0.35, 0.55, (300, 3)), xyz[:50] + 0.001]).astype(np.float32)
    st = G.stencil_cand(q)
    c27 = G.cand27(q)
    inv = np.float32(1.0) / h
    keys = np.where(xyz >= 0, (inv * xyz).astype(np.int32), (inv * xyz).astype(np.int32) - 1)
    for i, p in enumerate(q):
        assert st[i] == G.neighbors_fine(p)[1]
        kp = np.where(p >= 0, (inv * p).astype(np.int32), (inv * p).astype(np.int32) - 1)
        inside = np.all(np.abs(keys - kp) <= 1, axis=1).sum()
        c = kp - np.array(G.g.origin)
        dims = np.array(G.g.dims)
        if np.all(c >= 1) and np.all(c <= dims - 2):
            assert c27[i] == inside, i
        else:
            assert c27[i] == 0 and inside == 0, i
