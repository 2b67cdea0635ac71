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


def _densitygrid_brute(xyz, h, bmin, dims, W0):
    """Per node of the density grid: the particles strictly inside the L-infinity
    box of half-width h around the node centre (all n particles tested), and the
    flag of Dataset.cpp:136-163 with its 27-node weighted sum written out --
    weight exp(-1000 r) in FP32, r the node offset's L1 length."""
    cw = np.float32(h)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.stack([x.ravel(), y.ravel(), z.ravel()], -1).astype(np.float32)
    ctr = (bmin.astype(np.float32) + (idx + np.float32(0.5)) * cw).astype(np.float32)
    cnt = np.zeros(len(ctr), np.int64)
    cons = np.zeros(len(ctr), np.int64)
    p = xyz.astype(np.float32)
    R2 = np.float32(np.float32(1.5) * np.float32(h)) * np.float32(1.001)
    for s in range(0, len(ctr), 512):
        d = np.abs(p[None, :, :] - ctr[s:s + 512, None, :])
        cnt[s:s + 512] = np.all(d < cw, axis=2).sum(1)
        cons[s:s + 512] = np.all(d < R2, axis=2).sum(1)
    grid = cnt.reshape(nz, ny, nx).astype(np.float32)
    Nc = np.zeros_like(grid)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                w = np.exp(np.float32(-1000.0) * np.float32(abs(dx) + abs(dy) + abs(dz)), dtype=np.float32)
                sh = np.zeros_like(grid)
                zs, ys, xs = (slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)),
                              slice(max(0, -dx), nx - max(0, dx)))
                zd, yd, xd = (slice(max(0, dz), nz - max(0, -dz)), slice(max(0, dy), ny - max(0, -dy)),
                              slice(max(0, dx), nx - max(0, -dx)))
                sh[zs, ys, xs] = grid[zd, yd, xd]
                Nc = (Nc + w * sh).astype(np.float32)
    flag = (Nc.ravel() * np.float32(W0)) > np.float32(1.0)
    return ctr, cnt, flag, cons > 0


@pytest.mark.parametrize("seed,h", [(0, 0.04), (1, 2.0)])
def test_densitygrid_counts_equal_brute_force(seed, h):
    """Dataset.cpp:119-130 (per-node particle count from the box query around
    the node centre) and 143-163 (flag = approximate density > 1): the oracle's
    count, flag and conservative flag equal a brute force over every particle
    for every node, including particles placed exactly on box faces (excluded:
    the box test is strict) and nodes on the grid border.  With h = 0.04,
    W(0) = 8 / (pi h^3) > 1 and every occupied node is flagged; with h = 2,
    W(0) = 0.32 and the flag needs four particles, so the threshold matters."""
    rng = np.random.default_rng(seed)
    L = 15.0 * h
    xyz = np.concatenate([rng.uniform(0.0, L, (1500, 3)),
                          rng.normal(0.5 * L, 0.25 * h, (600, 3)),
                          rng.uniform(0.0, L, (4, 3)) * [1, 0, 0]]).astype(np.float32)
    xyz[-4:, 1:] = 0.0                                              # on the lower faces of the bounds
    G0 = O.Grid(xyz, h)
    dims = tuple(G0.densitygrid()[0])
    # particles exactly on faces of node boxes: centre +- h on one axis, kept
    # where the FP32 distance to the centre is exactly h
    m = 400
    node = rng.integers(2, np.array(dims) - 2, (m, 3)).astype(np.float32)
    ctr = (G0.bmin.astype(np.float32) + (node + np.float32(0.5)) * np.float32(h)).astype(np.float32)
    axis = rng.integers(0, 3, m)
    sign = rng.choice([-1.0, 1.0], m).astype(np.float32)
    face = ctr.copy()
    face[np.arange(m), axis] = (ctr[np.arange(m), axis] + sign * np.float32(h)).astype(np.float32)
    exact = np.abs(face - ctr).max(1) == np.float32(h)
    assert exact.sum() >= 40
    face, node, ctr = face[exact][:40], node[exact][:40], ctr[exact][:40]
    xyz = np.concatenate([xyz, face]).astype(np.float32)
    G = O.Grid(xyz, h)
    assert np.array_equal(G.bmin, G0.bmin)
    dims, count_box, flag_box, flag_cons = G.densitygrid()
    W0 = O.W(0.0, h)
    bctr, bcnt, bflag, bcons = _densitygrid_brute(xyz, h, G.bmin, dims, W0)
    lin = (node[:, 0] + dims[0] * (node[:, 1] + dims[1] * node[:, 2])).astype(np.int64)
    assert np.array_equal(bctr[lin], ctr)                           # the nodes the face particles touch
    assert np.array_equal(count_box.astype(np.int64), bcnt)
    assert np.array_equal(flag_box.astype(bool), bflag)
    assert np.array_equal(flag_cons.astype(bool), bcons)
    assert 0 < bflag.sum() < len(bflag) and np.any(bcnt == 0)
    if W0 < 1.0:
        assert np.any((bcnt > 0) & ~bflag)                          # occupied but below the threshold
    # the 26 neighbour weights vanish in FP32, so the flag is the node's own count
    assert np.exp(np.float32(-1000.0), dtype=np.float32) == 0.0
    assert np.array_equal(bflag, bcnt.astype(np.float32) * np.float32(W0) > np.float32(1.0))
