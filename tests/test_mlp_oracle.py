"""Pins of the step-proposal network oracle (oracle/mlp_oracle.py): torch float64
autograd golden vectors, finite differences, and the size of the BF16 effects --
no GPU."""
import pathlib

import numpy as np
import pytest

import features_np
import mlp_oracle as M
import oracle as O
from pmarch import scenes

GOLD = pathlib.Path(__file__).resolve().parent / "golden" / "mlp_fp64.npz"


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)        # plain arrays only (allow_pickle stays False)


def test_exact_backward_matches_autograd(gold):
    loss, g = M.loss_grad_exact(gold["p"], gold["x"], gold["t"], 4.0)
    assert loss == pytest.approx(float(gold["loss"]), rel=1e-12)
    assert np.allclose(g, gold["grad"], rtol=1e-9, atol=1e-12 * np.abs(gold["grad"]).max())


def test_exact_backward_matches_finite_differences(gold):
    p = gold["p"].astype(np.float64)
    x, t = gold["x"], gold["t"]
    _, g = M.loss_grad_exact(p, x, t)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(M.OFF_B1, 12, replace=False),                # W1 (incl. padding)
                          M.OFF_B1 + rng.choice(M.H, 4, replace=False),
                          M.OFF_W2 + rng.choice(M.H * M.H, 12, replace=False),
                          M.OFF_B2 + rng.choice(M.H, 4, replace=False),
                          M.OFF_W3 + rng.choice(M.H, 4, replace=False), [M.OFF_B3]])
    for i in idx:
        col = (i - M.OFF_W1) % M.K1 if i < M.OFF_B1 else 0
        if i < M.OFF_B1 and col >= M.NIN:
            assert g[i] == 0.0
            continue
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (M.loss_grad_exact(pp, x, t)[0] - M.loss_grad_exact(pm, x, t)[0]) / (2 * e)
        assert fd == pytest.approx(g[i], rel=2e-5, abs=1e-7), i


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.0048828125, -3.14159, 0.0, 65504.0], np.float32)
    b = M.bf16(v)
    assert b[0] == 1.0 and b[1] == 1.0          # tie to even
    assert b[2] == 1.0078125                     # above the tie rounds up
    assert abs(b[3] + 3.140625) < 1e-12
    assert b[4] == 0.0


def test_bf16_oracle_close_to_exact(gold):
    l1, g1 = M.loss_grad(gold["p"], gold["x"], gold["t"])
    l2, g2 = M.loss_grad_exact(gold["p"], gold["x"], gold["t"])
    assert l1 == pytest.approx(l2, rel=2e-2)
    cos = float(g1 @ g2 / (np.linalg.norm(g1) * np.linalg.norm(g2)))
    assert cos > 0.999


def test_init_layout():
    p = M.init(3)
    assert p.size == M.NPARAM == 6337
    W1 = p[M.OFF_W1:M.OFF_B1].reshape(M.H, M.K1)
    assert np.all(W1[:, M.NIN:] == 0) and np.any(W1[:, :M.NIN] != 0)
    # trainable parameters of the 24 -> 64 -> 64 -> 1 network
    assert M.NIN * M.H + M.H + M.H * M.H + M.H + M.H + 1 == 5825


def test_feature_restatement_gradient_matches_differences():
    """The numpy gradient of the features is the derivative of the numpy
    density: central differences in FP64 at random points of a lattice."""
    sc = scenes.lattice_cube(2, n_side=10)
    G = O.Grid(sc.xyz, sc.h)
    F = features_np.Features(sc.xyz, sc.h, sc.iso, G.origin, G.dims)
    rng = np.random.default_rng(4)
    for q in rng.uniform(0.35, 0.65, (50, 3)):
        rho, g, _ = F.rho_grad(q)
        assert rho == pytest.approx(G.density64(q[None])[0], rel=1e-12)
        e = 1e-6
        fd = [(F.rho_grad(q + e * u)[0] - F.rho_grad(q - e * u)[0]) / (2 * e) for u in np.eye(3)]
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * sc.iso / sc.h)
