#!/usr/bin/env python3
"""Throughput of the particle-fluid iso-surface march on one MI355X.

Workload (one "step" = one march of one view): the dam-break scene, 1,024,000
particles (80x160x80 lattice at 5 mm, r = 2.5 mm, h = 1 cm, Gaussian jitter
sigma 2 mm, seed 0), one 1920x1080 view from an orbit camera (30 deg
elevation, 45 deg azimuth, radius 1.3, aimed at the column centre), configured march: AABB entry, step
r/2, 8 bisection iterations (last in FP64), central-difference normals,
compensated FP32 density sums.  Particles and grid are resident in HBM before
the timed region; outputs stay in device memory.

Prints one JSON line (see the bench contract in DESIGN.md).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import platform
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "bachelor-thesis_amd"))

HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def alg_bytes(st: dict) -> float:
    """12 B per stencil candidate, 72 B of row bounds per sample, 32 B of output per ray."""
    return 12.0 * st["cand"] + 72.0 * st["samples"] + 32.0 * st["rays"]


def run_gpu(m, frame, settings, cam, steps, warmup):
    W, H = cam.width, cam.height
    for _ in range(warmup):
        m.march_async(frame, settings, cam.inv_pv, cam.eye, W, H, outputs=False)
        m.wait()
    kms = []
    stats = None
    t0 = time.perf_counter()
    for _ in range(steps):
        m.march_async(frame, settings, cam.inv_pv, cam.eye, W, H, outputs=False)
        m.wait()
        stats = m.stats()
        kms.append(stats["kernel_ms"])
    wall = time.perf_counter() - t0
    return wall, float(np.mean(kms)), stats


def cpu_baseline(scene, cam, settings, budget_s: float):
    """The oracle's reference-style marcher (per-sample hashed 27-cell query, one
    pixel per atomic fetch_add, hardware threads - 1) on a row subsample."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    threads = max(1, min(16, os.cpu_count() or 2) - 1)
    G = O.Grid(scene.xyz, scene.h)
    hs = O.HashedSearch(scene.xyz, scene.h)
    st = O.make_settings(**settings)
    W, H = cam.width, cam.height
    probe = 64
    t0 = time.perf_counter()
    hs.march(G, st, cam.inv_pv, cam.eye, W, H, row_step=probe, threads=threads)
    tp = time.perf_counter() - t0
    rows_probe = (H + probe - 1) // probe
    per_row = tp / rows_probe
    rows = max(rows_probe, min(H, int(budget_s / max(per_row, 1e-6))))
    step = max(1, H // rows)
    t0 = time.perf_counter()
    _, _, _, cnt = hs.march(G, st, cam.inv_pv, cam.eye, W, H, row_step=step, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": cnt["rays"] / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"every {step}th row of the same 1920x1080 view ({cnt['rays']} rays), "
                      f"{dt:.1f} s wall, {threads} threads on {os.cpu_count()} cpus ({cpu_model()})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import pmarch
    from pmarch import scenes

    scene = scenes.dam_break(0)
    cam = scene.cameras[0]
    plain = scene.settings()
    m = pmarch.Marcher(0)
    fi = m.set_frame(0, scene.xyz, scene.h)

    wall, kms, st = run_gpu(m, 0, plain, cam, args.steps, args.warmup)
    rays = cam.width * cam.height
    value = args.steps * rays / wall / 1e6
    B = alg_bytes(st)
    achieved = B / (kms * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s @1080p/1M particles, plain vs learned-step march, at fixed depth error",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "dam_break 1,024,000 particles, 1920x1080, plain march (AABB entry, step r/2, "
                               "8 bisections, last FP64, central-difference normals)",
                   "particles": scene.n, "width": cam.width, "height": cam.height, "h": scene.h,
                   "step": scene.step, "iso": scene.iso, "path": "plain"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "kernel_ms": round(kms, 4),
        "grid_build_ms": round(float(fi.build_ms), 3),
        "march_counters": {k: st[k] for k in ("rays", "rays_in_aabb", "hits", "samples", "cand", "march_samples")},
        "alg_bytes_per_launch": B,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, cam, plain, args.cpu_budget)
    m.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
