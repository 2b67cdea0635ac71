"""GPU parity of the step-proposal network (BF16 MFMA forward/backward, FP32
accumulation) against the FP64 oracle with the same BF16 roundings, of the
device training loop, and of the trace logger's rows against the CPU oracle."""
import math
import pathlib

import numpy as np
import pytest

import features_np
import mlp_oracle as M
import oracle as O
from features_np import point
from pmarch import scenes

pytestmark = pytest.mark.gpu
GOLD = pathlib.Path(__file__).resolve().parent / "golden" / "mlp_fp64.npz"


def test_mlp_forward_matches_oracle(marcher):
    rng = np.random.default_rng(0)
    p = M.init(11)
    p[M.OFF_B1:M.OFF_W2] = rng.uniform(-0.3, 0.3, M.H)
    p[M.OFF_B2:M.OFF_W3] = rng.uniform(-0.3, 0.3, M.H)
    marcher.mlp_set(p)
    assert np.array_equal(marcher.mlp_get(), p)
    x = rng.uniform(-1, 2, (1000, M.NIN)).astype(np.float32)     # ragged: not a multiple of 64
    y = marcher.mlp_forward(x)
    yo, _ = M.forward(p, x)
    err = np.abs(y - yo)
    print("forward max err", err.max(), "max |y|", np.abs(yo).max(), flush=True)
    assert np.all(err <= 2e-3 * (1.0 + np.abs(yo)))


def test_mlp_gradient_matches_oracle(marcher):
    g = np.load(GOLD)
    marcher.mlp_set(g["p"])
    loss, grad = marcher.mlp_grad(g["x"], g["t"], 4.0)
    lo, go = M.loss_grad(g["p"], g["x"], g["t"], 4.0)
    assert loss == pytest.approx(lo, rel=2e-3)
    for name, a, b in (("W1", M.OFF_W1, M.OFF_B1), ("b1", M.OFF_B1, M.OFF_W2), ("W2", M.OFF_W2, M.OFF_B2),
                       ("b2", M.OFF_B2, M.OFF_W3), ("w3", M.OFF_W3, M.OFF_B3), ("b3", M.OFF_B3, M.NPARAM)):
        u, v = grad[a:b].astype(np.float64), go[a:b]
        rel = np.linalg.norm(u - v) / max(np.linalg.norm(v), 1e-30)
        print(name, "relative error", rel, flush=True)
        assert rel < 2e-2, name
    # and against float64 autograd of the unrounded network
    cos = float(grad @ g["grad"] / (np.linalg.norm(grad) * np.linalg.norm(g["grad"])))
    assert cos > 0.999
    assert np.all(grad[M.OFF_W1:M.OFF_B1].reshape(M.H, M.K1)[:, M.NIN:] == 0)


def test_mlp_training_fits_smooth_target(marcher):
    """Adam on a device dataset of traces lowers the loss; determinism of two runs."""
    sc = scenes.dam_break(5, nx=40, ny=80, nz=40, width=480, height=270)
    c = sc.cameras[0]
    marcher.set_frame(20, sc.xyz, sc.h)
    marcher.dataset_clear(200_000)
    n = marcher.trace_append(20, sc.settings(), c.inv_pv, c.eye, c.width, c.height, ray_stride=4, k_stride=2, seed=1)
    assert n > 10_000
    losses = []
    params = []
    for _ in range(2):
        marcher.mlp_set(M.init(0))
        losses.append(marcher.mlp_train(4, batch=1024, lr=2e-3, w_over=4.0, seed=5))
        params.append(marcher.mlp_get())
    print("loss per epoch", losses[0], flush=True)
    assert losses[0][-1] < 0.5 * losses[0][0]
    assert np.array_equal(params[0], params[1]) and np.array_equal(losses[0], losses[1])


def test_trace_rows_match_oracle(marcher):
    sc = scenes.lattice_cube(2, width=160, height=120)
    c = sc.cameras[0]
    marcher.set_frame(21, sc.xyz, sc.h)
    marcher.dataset_clear(50_000)
    n = marcher.trace_append(21, sc.settings(), c.inv_pv, c.eye, c.width, c.height, ray_stride=3, k_stride=2, seed=9)
    feat, label, hit, meta = marcher.dataset_get(0, n)
    assert n > 2000 and np.all(np.isfinite(feat)) and np.all(label >= 0)
    G = O.Grid(sc.xyz, sc.h)
    _, _, t, _ = O.march(G, O.make_settings(**sc.settings()), c.inv_pv, c.eye, c.width, c.height)
    F = features_np.Features(sc.xyz, sc.h, sc.iso, G.origin, G.dims)
    rng = np.random.default_rng(0)
    rows = rng.choice(n, 300, replace=False)
    varied = np.zeros(4, np.int64)
    for i in rows:
        pix, k = int(meta[i, 0]), int(meta[i, 1])
        x, y = pix % c.width, pix // c.width
        o, d = O.ray(c.inv_pv, c.eye, c.width, c.height, x, y)
        ok, tn, tf = O.slab(o, d, G.bmin, G.bmax)
        assert ok
        tk = np.float32(tn) + np.float32(k) * np.float32(sc.step)
        th = t[y, x]
        assert bool(hit[i]) == (th >= 0)
        tend = th if th >= 0 else tf
        expect = math.log2(1.0 + min((tend - float(tk)) / sc.h, 255.0))
        assert abs(label[i] - expect) < 1e-4, (i, label[i], expect)
        # all 24 features against the numpy restatement: counts exactly, the
        # density to FP32 rounding, |grad rho| (plain FP32 sums) within a bound
        # from the sum of the terms' magnitudes
        f = F.row(o, d, tk, tn, tf)
        counts = np.r_[0:16, 23]
        assert np.array_equal(feat[i, counts], f[counts].astype(np.float32)), (i, feat[i, counts], f[counts])
        assert np.allclose(feat[i, 16:19], f[16:19], rtol=0, atol=1e-6)
        assert feat[i, 19] == pytest.approx(f[19], rel=1e-5, abs=1e-6)
        assert feat[i, 22] == pytest.approx(f[22], rel=1e-5, abs=1e-6)
        assert feat[i, 20] == pytest.approx(f[20], rel=5e-6, abs=2e-7)
        _, g, scale = F.rho_grad(point(o, tk, d))
        tol = 1e-5 * np.linalg.norm(scale) * sc.h / sc.iso / 4.0
        assert abs(feat[i, 21] - f[21]) <= tol + 1e-6, (i, feat[i, 21], f[21], tol)
        varied += feat[i, [0, 8, 21, 23]] != feat[rows[0], [0, 8, 21, 23]]
    assert np.all(varied > 0)
