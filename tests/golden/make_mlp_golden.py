"""Writes tests/golden/mlp_fp64.npz: a seeded batch, parameters, and the loss and
gradient of the step-proposal network computed by torch float64 autograd -- an
implementation independent of the oracle's hand-written backward pass.

Run from the repository root:  python tests/golden/make_mlp_golden.py
"""
import pathlib
import sys

import numpy as np
import torch

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as M  # noqa: E402


def main():
    rng = np.random.default_rng(1234)
    n = 128
    x = rng.uniform(-1.0, 2.0, (n, M.NIN)).astype(np.float32)
    t = rng.uniform(0.0, 6.0, n).astype(np.float32)
    p = M.init(7)
    p[M.OFF_B1:M.OFF_W2] = rng.uniform(-0.2, 0.2, M.H)
    p[M.OFF_B2:M.OFF_W3] = rng.uniform(-0.2, 0.2, M.H)
    P = torch.tensor(p.astype(np.float64), requires_grad=True)
    W1 = P[M.OFF_W1:M.OFF_B1].reshape(M.H, M.K1)
    b1 = P[M.OFF_B1:M.OFF_W2]
    W2 = P[M.OFF_W2:M.OFF_B2].reshape(M.H, M.H)
    b2 = P[M.OFF_B2:M.OFF_W3]
    w3 = P[M.OFF_W3:M.OFF_B3]
    b3 = P[M.OFF_B3]
    X = torch.zeros(n, M.K1, dtype=torch.float64)
    X[:, :M.NIN] = torch.tensor(x.astype(np.float64))

    def act(z):
        return z * torch.sigmoid(1.702 * z)

    y = act(act(X @ W1.T + b1) @ W2.T + b2) @ w3 + b3
    d = y - torch.tensor(t.astype(np.float64))
    w = torch.where(d > 0, torch.full_like(d, 4.0), torch.ones_like(d))
    loss = (w * d * d).sum() / n
    loss.backward()
    g = P.grad.detach().numpy().copy()
    g[M.OFF_W1:M.OFF_B1].reshape(M.H, M.K1)[:, M.NIN:] = 0.0
    out = ROOT / "tests" / "golden" / "mlp_fp64.npz"
    np.savez_compressed(out, x=x, t=t, p=p, y=y.detach().numpy(), loss=np.float64(loss.item()), grad=g)
    print("wrote", out, "loss", loss.item())


if __name__ == "__main__":
    main()
