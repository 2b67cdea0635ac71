"""Pins of the oracle's SPH kernel and density (Kernel.cpp:8-52, RayMarcher.cpp:321-325)
against closed forms and quadrature -- no GPU."""
import math

import numpy as np
import pytest

import oracle as O


def closed_W(q):
    """Standard 3-D cubic spline shape (sigma factored out), FP64."""
    if q < 0.5:
        return 6.0 * (q ** 3 - q ** 2) + 1.0
    if q < 1.0:
        return 2.0 * (1.0 - q) ** 3
    return 0.0


def closed_lattice_sum():
    """sum over the lattice of spacing h/2 of W/sigma: shells at |n| = 1, sqrt2, sqrt3, 2."""
    return (1.0 + 6 * closed_W(0.5) + 12 * closed_W(math.sqrt(2) / 2) + 8 * closed_W(math.sqrt(3) / 2)
            + 6 * closed_W(1.0))


def test_sigma_and_special_values():
    h = 0.125  # exact in binary: (h/2)^2, sqrt and *1/h are exact
    sig = O.sigma(h)
    assert sig == np.float32(8.0) / (np.float32(math.pi) * np.float32(h) ** 3)
    assert O.W(0.0, h) == sig                       # q = 0
    assert O.W((h / 2) ** 2, h) == np.float32(sig * 0.25)   # q = 1/2 from the outer branch
    assert O.W(h * h, h) == 0.0                     # q = 1, strict support
    assert O.W(h * h * 1.5, h) == 0.0
    for q in (0.1, 0.25, 0.4, 0.6, 0.75, 0.9):
        r2 = (q * h) ** 2
        assert O.W64(r2, h) == pytest.approx(8 / (math.pi * h ** 3) * closed_W(q), rel=1e-13)
        assert O.W(r2, h) == pytest.approx(sig * closed_W(q), rel=3e-6)


def test_continuity_at_half():
    h = 0.1
    eps = 1e-9
    lo = O.W64(((0.5 - eps) * h) ** 2, h)
    hi = O.W64(((0.5 + eps) * h) ** 2, h)
    assert abs(lo - hi) / O.W64(0, h) < 1e-7


@pytest.mark.parametrize("h", [0.1, 0.01, 0.012])
def test_normalised_by_quadrature(h):
    """int W dV = 1 (Simpson on 4 pi r^2 W(r)); catches a wrong sigma or branch polynomial."""
    n = 20001
    r = np.linspace(0.0, h, n)
    f = np.array([O.W64(x * x, h) for x in r]) * 4 * math.pi * r * r
    w = np.ones(n)
    w[1:-1:2] = 4
    w[2:-1:2] = 2
    integral = (r[1] - r[0]) / 3 * (f * w).sum()
    assert integral == pytest.approx(1.0, abs=1e-9)


def test_gradW_has_extra_inverse_distance():
    """gradW(r) = dW/d|r| * r_hat / |r| (Kernel.cpp:43 divides normalize(r) by |r| h)."""
    h = 0.1
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        q = rng.uniform(0.05, 0.95)
        r = (v * q * h).astype(np.float32)
        rl = float(np.linalg.norm(r.astype(np.float64)))
        dq = 1e-6
        dWdr = (O.W64(((q + dq) * h) ** 2, h) - O.W64(((q - dq) * h) ** 2, h)) / (2 * dq * h)
        expect = dWdr * (r.astype(np.float64) / rl) / rl
        got = O.gradW(r, h).astype(np.float64)
        assert np.allclose(got, expect, rtol=2e-4, atol=1e-3 * np.abs(expect).max()), (q, got, expect)
    # zero outside the support, and outward (away from the particle) for an outer sample
    assert np.all(O.gradW(np.array([h, 0, 0], np.float32), h) == 0)
    g = O.gradW(np.array([0.7 * h, 0, 0], np.float32), h)
    assert g[0] < 0 and g[1] == 0 and g[2] == 0


def test_perfect_lattice_density():
    """rho at a site of a perfect lattice of spacing h/2 = sigma * 3.1415061532636885."""
    assert closed_lattice_sum() == pytest.approx(O.RESTING_LATTICE_SUM, rel=1e-15)
    h = 0.125
    s = h / 2
    i = np.arange(-5, 6)
    g = (np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3) * s + 0.3125).astype(np.float32)
    G = O.Grid(g, h)
    site = np.array([[0.3125, 0.3125, 0.3125]])
    sig_d = 8 / (math.pi * h ** 3)
    assert G.density64(site)[0] / sig_d == pytest.approx(O.RESTING_LATTICE_SUM, rel=1e-14)
    r32 = G.density(site.astype(np.float32))[0]
    assert r32 / O.sigma(h) == pytest.approx(O.RESTING_LATTICE_SUM, rel=2e-7)
    # every interior site sees exactly the 27 non-zero terms
    _, cand = G.density(g, with_cand=True)
    nb, _ = G.neighbors(site[0].astype(np.float32))
    assert len(nb) == 27


def test_compensated_sum_is_closer_to_fp64():
    sys_rng = np.random.default_rng(4)
    xyz = sys_rng.uniform(0, 0.2, (40000, 3)).astype(np.float32)   # ~200 neighbours per point at h=0.05
    G = O.Grid(xyz, 0.05)
    q = sys_rng.uniform(0.05, 0.15, (2000, 3)).astype(np.float32)
    r64 = G.density64(q.astype(np.float64))
    ec = np.abs(G.density(q, True).astype(np.float64) - r64) / r64
    ep = np.abs(G.density(q, False).astype(np.float64) - r64) / r64
    assert ec.max() <= 4e-7            # a few ulp: each FP32 term carries its own rounding
    assert ec.mean() < ep.mean()
