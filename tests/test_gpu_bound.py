"""GPU sub-box bound skip: the bits against the oracle's, and the march with
the skip against the oracle march without it (bit for bit) and the oracle
march with it (same counters)."""
import numpy as np
import pytest

import oracle as O
from pmarch import scenes

pytestmark = pytest.mark.gpu

KEYS = dict(O.Settings._fields_)


def _oracle_settings(sc, **kw):
    return O.make_settings(**{k: v for k, v in sc.settings(**kw).items() if k in KEYS})


@pytest.mark.parametrize("boxes", [0, 1])
@pytest.mark.parametrize("name", ["lattice", "dam_small", "clustered"])
def test_bound_bits_match_oracle(marcher, name, boxes):
    sc = {"lattice": lambda: scenes.lattice_cube(2),
          "dam_small": lambda: scenes.dam_break(4, nx=30, ny=60, nz=30),
          "clustered": lambda: scenes.clustered_pool(3, n_total=30000, r=0.008, width=96, height=64)}[name]()
    print(f"bound bits {name} n={sc.n} boxes={boxes}", flush=True)
    marcher.set_frame(40, sc.xyz, sc.h)
    gpu = marcher.bound_grid(40, sc.iso, boxes)
    ref = O.Grid(sc.xyz, sc.h).bound(sc.iso, boxes)
    assert gpu.shape == ref.shape
    assert np.array_equal(gpu, ref), f"{np.sum(gpu != ref)} words differ"
    if boxes == 1:
        b = gpu.view(np.uint8)
        assert np.all((b == 0) | (b == 0xFF))


@pytest.mark.parametrize("boxes", [0, 1])
@pytest.mark.parametrize("name", ["lattice", "clustered"])
def test_bound_march_matches_oracle(marcher, name, boxes):
    sc = {"lattice": lambda: scenes.lattice_cube(1, width=200, height=150),
          "clustered": lambda: scenes.clustered_pool(2, n_total=40000, r=0.008, width=200, height=150)}[name]()
    c = sc.cameras[0]
    print(f"bound march {name} n={sc.n} {c.width}x{c.height} boxes={boxes}", flush=True)
    marcher.set_frame(41, sc.xyz, sc.h)
    pos, nrm, t, st = marcher.march(41, sc.settings(skip_mode=4, bound_boxes=boxes), c.inv_pv, c.eye, c.width,
                                    c.height)
    G = O.Grid(sc.xyz, sc.h)
    p0, n0, t0, c0 = O.march(G, _oracle_settings(sc, skip_mode=0), c.inv_pv, c.eye, c.width, c.height)
    assert np.array_equal(pos, p0) and np.array_equal(nrm, n0) and np.array_equal(t, t0)
    p4, n4, t4, c4 = O.march(G, _oracle_settings(sc, skip_mode=4, bound_boxes=boxes), c.inv_pv, c.eye, c.width,
                             c.height)
    for k in ("rays", "rays_in_aabb", "hits", "samples", "march_cand", "march_samples"):
        assert st[k] == c4[k], (k, st[k], c4[k])
    print("march samples", c0["march_samples"], "->", st["march_samples"], flush=True)
    assert st["march_samples"] < c0["march_samples"]


def test_bound_march_dam_break_1080p_identical(marcher):
    """Full size: the skip changes no pixel of the 1080p dam-break view."""
    sc = scenes.dam_break(0)
    c = sc.cameras[0]
    print(f"bound 1080p dam_break n={sc.n}", flush=True)
    marcher.set_frame(42, sc.xyz, sc.h)
    a = marcher.march(42, sc.settings(skip_mode=0), c.inv_pv, c.eye, c.width, c.height)
    b = marcher.march(42, sc.settings(skip_mode=4), c.inv_pv, c.eye, c.width, c.height)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    print("samples", a[3]["samples"], "->", b[3]["samples"], "kernel ms", a[3]["kernel_ms"], "->",
          b[3]["kernel_ms"], flush=True)
    assert b[3]["samples"] < a[3]["samples"]


def test_bound_rejected_in_reference_mode(marcher):
    import pmarch
    sc = scenes.lattice_cube(0, n_side=8, width=32, height=24)
    marcher.set_frame(43, sc.xyz, sc.h)
    c = sc.cameras[0]
    with pytest.raises(pmarch.PmError):
        marcher.march(43, sc.settings(mode=1, skip_mode=4, max_steps=100), c.inv_pv, c.eye, 32, 24,
                      depth=np.ones((24, 32), np.float32))


def test_bound_boxes_invalid_rejected(marcher):
    """bound_boxes other than 0 or 1 is refused before any launch, by the march
    and by pm_get_bound_grid."""
    import pmarch
    sc = scenes.lattice_cube(0, n_side=8, width=32, height=24)
    marcher.set_frame(44, sc.xyz, sc.h)
    c = sc.cameras[0]
    for bad in (2, -1):
        with pytest.raises(pmarch.PmError):
            marcher.march(44, sc.settings(skip_mode=4, bound_boxes=bad), c.inv_pv, c.eye, 32, 24)
        with pytest.raises(pmarch.PmError):
            marcher.bound_grid(44, sc.iso, bad)
    # a valid march still runs afterwards
    pos = marcher.march(44, sc.settings(skip_mode=4, bound_boxes=1), c.inv_pv, c.eye, 32, 24)[0]
    assert pos.shape == (24, 32, 4)


def test_bound_boxes_switch_rebuilds(marcher):
    """The bits are keyed by iso and bound_boxes: marching one frame with 8x8x8,
    then 4x4x4, then 8x8x8 boxes gives the oracle's counters for each, and the
    same image every time."""
    sc = scenes.clustered_pool(2, n_total=20000, r=0.008, width=96, height=64)
    c = sc.cameras[0]
    marcher.set_frame(45, sc.xyz, sc.h)
    G = O.Grid(sc.xyz, sc.h)
    ref = {}
    for boxes in (0, 1):
        ref[boxes] = O.march(G, _oracle_settings(sc, skip_mode=4, bound_boxes=boxes), c.inv_pv, c.eye, c.width,
                             c.height)
    first = None
    for boxes in (0, 1, 0):
        out = marcher.march(45, sc.settings(skip_mode=4, bound_boxes=boxes), c.inv_pv, c.eye, c.width, c.height)
        for k in ("samples", "march_samples", "hits"):
            assert out[3][k] == ref[boxes][3][k], (boxes, k)
        if first is None:
            first = out
        for a, b in zip(first[:3], out[:3]):
            assert np.array_equal(a, b)
    assert ref[1][3]["march_samples"] > ref[0][3]["march_samples"]
