"""Training of the step-proposal network from plain-march traces, on the GPU.

The plain configured march is the teacher: for sampled rays of seeded scenes
(dam-break columns and clustered pools with random orbit cameras) the trace
logger writes, for every few march samples before the crossing, the network's
24 inputs and the label log2(1 + d/h), d the distance along the ray to the
first iso crossing (or to the AABB exit on a miss).  Training runs Adam on the
device (forward/backward with BF16 MFMA, FP32 accumulation).

    python -m pmarch.train --out weights/step_mlp.npz      (from bachelor-thesis_amd/)

Checkpoint: an .npz of plain arrays (no pickle): params (float32, 6337),
version, feature_version, and the training summary.
"""
from __future__ import annotations

import argparse
import json
import pathlib
import time

import numpy as np

from . import scenes
from ._lib import MLP_PARAMS, Marcher

CHECKPOINT_VERSION = 1
FEATURE_VERSION = 1
WEIGHTS = pathlib.Path(__file__).resolve().parent / "weights" / "step_mlp.npz"

W_OVER = 4.0


def he_init(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases, output bias log2(9); padding inputs zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, np.float32)
    W1 = np.zeros((64, 32), np.float32)
    W1[:, :24] = rng.uniform(-1, 1, (64, 24)) * np.sqrt(6.0 / 24)
    p[0:2048] = W1.reshape(-1)
    p[2048 + 64:2048 + 64 + 4096] = (rng.uniform(-1, 1, (64, 64)) * np.sqrt(6.0 / 64)).reshape(-1)
    o_w3 = 2048 + 64 + 4096 + 64
    p[o_w3:o_w3 + 64] = rng.uniform(-1, 1, 64) * np.sqrt(3.0 / 64)
    p[o_w3 + 64] = np.log2(9.0)
    return p


def training_views(n_scenes: int, seed0: int, width: int, height: int):
    """Yields (scene, camera) pairs: n_scenes dam-break columns and n_scenes
    clustered pools (1 M particles each), one random orbit camera per scene."""
    for i in range(n_scenes):
        rng = np.random.default_rng(seed0 + i)
        s = scenes.dam_break(seed0 + i, width=width, height=height)
        cam = scenes.random_views(rng, scenes.COLUMN_CENTRE, 1, width, width, height, height, radius=(1.1, 1.9))[0]
        yield s, cam
        p = scenes.clustered_pool(seed0 + 1000 + i, n_total=1_000_000, width=width, height=height)
        cam = scenes.random_views(rng, (0.5, 0.3, 0.2), 1, width, width, height, height, radius=(1.2, 2.0))[0]
        yield p, cam


def collect(m: Marcher, rows: int, n_scenes: int = 20, seed0: int = 100, width: int = 1280, height: int = 720,
            ray_stride: int = 4, k_stride: int = 3, log=print) -> int:
    """Equal row budgets per scene (rows / 2 n_scenes), so no single view dominates."""
    m.dataset_clear(rows)
    per_scene = rows // (2 * n_scenes)
    total = 0
    for j, (s, cam) in enumerate(training_views(n_scenes, seed0, width, height)):
        m.set_frame(900, s.xyz, s.h)
        total = m.trace_append(900, s.settings(), cam.inv_pv, cam.eye, cam.width, cam.height, ray_stride, k_stride,
                               seed=seed0 + j, max_new_rows=per_scene)
        log(f"trace {j:2d} {s.name:15s} rows {total}")
    return total


def save(path: pathlib.Path, params: np.ndarray, summary: dict):
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32), version=np.int32(CHECKPOINT_VERSION),
             feature_version=np.int32(FEATURE_VERSION), summary=np.frombuffer(json.dumps(summary).encode(), np.uint8))


def load(path: pathlib.Path = WEIGHTS) -> tuple[np.ndarray, dict]:
    z = np.load(path)                   # plain arrays, allow_pickle stays False
    if int(z["version"]) != CHECKPOINT_VERSION or int(z["feature_version"]) != FEATURE_VERSION:
        raise ValueError(f"checkpoint {path}: version mismatch")
    return z["params"].astype(np.float32), json.loads(bytes(z["summary"]).decode())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=2_000_000)
    ap.add_argument("--scenes", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", type=pathlib.Path, default=WEIGHTS)
    args = ap.parse_args()
    m = Marcher(0)
    t0 = time.time()
    n = collect(m, args.rows, args.scenes)
    t1 = time.time()
    m.mlp_set(he_init(args.seed))
    losses = m.mlp_train(args.epochs, args.batch, args.lr, W_OVER, args.seed)
    t2 = time.time()
    summary = dict(rows=n, scenes=2 * args.scenes, epochs=args.epochs, batch=args.batch, lr=args.lr, w_over=W_OVER,
                   loss_per_epoch=[float(v) for v in losses], trace_s=round(t1 - t0, 2), train_s=round(t2 - t1, 2))
    save(args.out, m.mlp_get(), summary)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
