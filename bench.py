#!/usr/bin/env python3
"""Throughput of the particle-fluid iso-surface march on one MI355X.

Default workload (one "step" = one march of one view): the dam-break scene,
1,024,000 particles (80x160x80 lattice at 5 mm, r = 2.5 mm, h = 1 cm, Gaussian
jitter sigma 2 mm, seed 0), one 1920x1080 view from an orbit camera (30 deg
elevation, 45 deg azimuth, radius 1.3, aimed at the column centre), configured
march: AABB entry, step r/2, 8 bisection iterations (last in FP64),
central-difference normals, compensated FP32 density sums, sub-box bound skip
(exact: the image is that of the march without it).  Particles and grid are
resident in HBM before the timed region; outputs stay in device memory.

Three forms are timed, with identical images (every pixel is checked): the
march with a fixed step from the AABB entry (the configuration as written), the
same with the exact sub-box bound skip, and the learned step proposal on top of
the bound skip (network from pmarch/weights/step_mlp.npz; its jumps are
certified, so its image is the plain one).  `value` is the fastest form's
Mrays/s; all three are reported, with the roofline of the value form from an
instrumented launch of the same view, the PCIe-inclusive rate, the bound-bit
build time, the CPU baseline and the FP64 errors of the image.  Prints one
JSON line.

--workload clustered_4k: the 4,000,000-particle clustered pool at 3840x2160,
same march and forms.

--workload batch: 64 views (W ~ U{256..1920}, H ~ U{256..1080}, seeded orbit
cameras) over 16 consecutive frames of the 1M dam break (1 % displacement per
frame), all 1,024 view-frames in one persistent launch replayed from a
hipGraph; one step = one replay of the whole batch.

--workload aniso_dam_1080p: the anisotropic march (PerPixel_Anisotropic, the
reference's default mode: per sample WPCA over the particles within 2h, then
the anisotropic kernel) on the dam-break column scaled to h = 0.1 (the
particle radius its k_s = 2000 is tuned for), 1920x1080, IsoDensity 1.0,
step h/8, 8 bisections (last in FP64), central-difference normals, exact skip of
the sub-boxes with no particle within h (the density is 0 there).

--workload autoplay_dam_1080p: the reference's autoplay, a new frame for every
march: 16 consecutive frames of the 1M dam break (1 % displacement per frame)
played through two frame slots, each frame's upload and grid build on the
second stream while the previous frame is marched (pm_set_frame_stream), and
the same frames played serially (set, then march) for comparison; one step =
one play of the 16 frames.

--workload sprites_dam_1080p: the depth-sprite pass (DepthRenderPass: every
particle of the 1M dam break as a sphere sprite of radius h, depth test LESS)
for the 1080p view into device memory; one step = one pass.

--workload reference_dam_1080p: the reference's own isotropic march (the
facade's path) from the GPU sprite pass's depths: MaxSteps 128, DensityGrid
box skip, normalize(sum gradW), plain sums.

--workload depth_smooth_1080p: the screen-space depth smoothing
(GaussRenderPass, N = 8: a 17x17 window) of a 1920x1080 depth image resident
in device memory; one step = one pass over the image.
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
TRAFFIC = [ROOT / "profiles" / "march_dam_break_1080p_traffic.json",
           ROOT / "profiles" / "march_aniso_dam_break_1080p_traffic.json",
           ROOT / "profiles" / "depth_smooth_1080p_traffic.json"]


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def alg_bytes(st: dict, cand: str = "cand_stencil") -> float:
    """Algorithmic bytes of a launch: 12 B (x, y, z) per candidate
    of every density evaluation, 72 B of stencil row bounds per evaluation and
    32 B of positions and normals per ray.  The candidates are counted per
    evaluation however the kernels took it (an instrumented launch of the same
    view counts them): cand_stencil -- the particles of the evaluation's own
    trimmed fine stencil, what a single-sample walk of this build reads -- or
    cand27 -- the particles of its 27 search cells, the reference's
    CompactNSearch query."""
    return 12.0 * st[cand] + 72.0 * st["samples"] + 32.0 * st["rays"]


def instrumented(m, frame, settings, cam) -> dict:
    """Counts of an instrumented (untimed) launch of the same march: every
    density evaluation's own-stencil and 27-cell candidates."""
    m.march_async(frame, dict(settings, instrument=1), cam.inv_pv, cam.eye, cam.width, cam.height, outputs=False)
    m.wait()
    return m.stats()


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


def image_errors(img, ref, ext: float, step: int) -> dict:
    """Errors of a march's image (rows 0, step, 2 step, ...) against a
    reference march (the FP64 oracle): mask agreement, max depth error (along
    the ray) over the scene extent, max normal angle."""
    pa, na, ta = (x[::step] for x in img)
    pb, nb, tb = (x[::step] for x in ref)
    ma, mb = pa[..., 3] > 0, pb[..., 3] > 0
    both = ma & mb
    dt = np.abs(ta[both].astype(np.float64) - tb[both].astype(np.float64))
    u, v = na[both][:, :3].astype(np.float64), nb[both][:, :3].astype(np.float64)
    ang = np.degrees(np.arctan2(np.linalg.norm(np.cross(u, v), axis=-1), (u * v).sum(-1)))
    return {"rows": int(pa.shape[0]), "mask_agree": round(float((ma == mb).mean()), 6),
            "depth_max_over_extent": float(dt.max() / ext) if dt.size else 0.0,
            "normal_max_deg": float(ang.max()) if ang.size else 0.0,
            "within_tolerance": bool(float((ma == mb).mean()) >= 0.999 and (dt.max() if dt.size else 0) <= 1e-4 * ext
                                     and (ang.max() if ang.size else 0) <= 0.5)}


def fp64_errors(scene, cam, img, step: int, settings=None, xyz=None) -> dict:
    """The oracle leg's check of the GPU image: the FP64 oracle march (every
    operation in double, all bisection steps in double) on every step-th row of
    the same view, against the GPU's rows."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    xyz = scene.xyz if xyz is None else xyz
    settings = scene.settings() if settings is None else settings
    G = O.Grid(xyz, scene.h)
    st = O.make_settings(**{k: v for k, v in settings.items() if k in dict(O.Settings._fields_)})
    ref = O.march64(G, st, cam.inv_pv, cam.eye, cam.width, cam.height, row_step=step)
    ext = float((xyz.max(0) - xyz.min(0)).max())
    return image_errors(img, ref, ext, step)


def merge_errors(errs) -> dict:
    """The worst of several images' errors."""
    return {"images": len(errs), "rows": sum(e["rows"] for e in errs),
            "mask_agree": min(e["mask_agree"] for e in errs),
            "depth_max_over_extent": max(e["depth_max_over_extent"] for e in errs),
            "normal_max_deg": max(e["normal_max_deg"] for e in errs),
            "within_tolerance": all(e["within_tolerance"] for e in errs)}


def cpu_baseline(scene, cam, settings, budget_s: float, xyz=None):
    """The oracle's reference-style marcher (per-sample hashed 27-cell query, one
    pixel per atomic fetch_add, hardware threads - 1) on the same view; a row
    subsample when a full frame would exceed the budget."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    threads = max(1, min(16, os.cpu_count() or 2) - 1)
    xyz = scene.xyz if xyz is None else xyz
    G = O.Grid(xyz, scene.h)
    hs = O.HashedSearch(xyz, scene.h)
    st = O.make_settings(**{k: v for k, v in settings.items() if k in dict(O.Settings._fields_)})
    W, H = cam.width, cam.height
    probe = 64
    t0 = time.perf_counter()
    hs.march(G, st, cam.inv_pv, cam.eye, W, H, row_step=probe, threads=threads)
    tp = time.perf_counter() - t0
    rows_probe = (H + probe - 1) // probe
    rows = max(rows_probe, min(H, int(budget_s / max(tp / rows_probe, 1e-6))))
    step = max(1, H // rows)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        _, _, _, cnt = hs.march(G, st, cam.inv_pv, cam.eye, W, H, row_step=step, threads=threads)
        runs.append(time.perf_counter() - t0)
        if sum(runs) > budget_s:
            break
    dt = float(np.median(runs))
    return {"value": round(cnt["rays"] / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{'every %dth row of ' % step if step > 1 else ''}the same {W}x{H} view "
                      f"({cnt['rays']} rays), median of {len(runs)} runs of {dt:.2f} s, {threads} threads on "
                      f"{os.cpu_count()} cpus ({cpu_model()})"}


def traffic_for(path: str, workload: str):
    """HBM bytes per launch of the profiled form of a workload (FETCH_SIZE +
    WRITE_SIZE of all its kernels, scaled against the copy kernel)."""
    for f in TRAFFIC:
        if f.exists():
            tr = json.loads(f.read_text())
            if tr.get("path") == path and tr.get("workload", "dam_break_1080p") == workload:
                return tr["hbm_bytes_per_launch"]
    return None


def alg_bytes_aniso(st: dict) -> float:
    """Anisotropic march: 12 B per candidate of its two walks (the 2h
    search-grid stencil for WPCA, the trimmed h march-grid stencil for the
    density: the candidate counter counts both), 8 B of row bounds per row
    (25 + 25 rows per sample) and 36 B of output per ray."""
    return 12.0 * st["cand"] + 8.0 * 50 * st["samples"] + 36.0 * st["rays"]


def cpu_baseline_aniso(scene, cam, settings, budget_s: float):
    """The oracle's anisotropic march (the same algorithm, grid walks, one pixel
    per atomic fetch_add) on a row subsample of the same view."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    threads = max(1, min(16, os.cpu_count() or 2) - 1)
    G = O.Grid(scene.xyz, scene.h)
    st = O.make_settings(**{k: v for k, v in settings.items() if k in dict(O.Settings._fields_)})
    if st.skip_mode == 4:
        G.bound(O.ZERO_ISO)        # the skip's bits, built once outside the timed runs
    W, H = cam.width, cam.height
    probe = 64
    t0 = time.perf_counter()
    O.march(G, st, cam.inv_pv, cam.eye, W, H, row_step=probe, threads=threads, want_t=False)
    tp = time.perf_counter() - t0
    rows_probe = (H + probe - 1) // probe
    rows = max(rows_probe, min(H, int(budget_s / max(tp / rows_probe, 1e-6))))
    step = max(1, H // rows)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        _, _, _, cnt = O.march(G, st, cam.inv_pv, cam.eye, W, H, row_step=step, threads=threads, want_t=False)
        runs.append(time.perf_counter() - t0)
        if sum(runs) > budget_s:
            break
    dt = float(np.median(runs))
    return {"value": round(cnt["rays"] / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{'every %dth row of ' % step if step > 1 else ''}the same {W}x{H} view "
                      f"({cnt['rays']} rays), median of {len(runs)} runs of {dt:.2f} s, {threads} threads on "
                      f"{os.cpu_count()} cpus ({cpu_model()})"}


def bench_aniso(args):
    """The anisotropic march on the dam-break column scaled to h = 0.1, 1080p."""
    import pmarch
    from pmarch import scenes

    scene = scenes.dam_break_aniso(0)
    cam = scene.cameras[0]
    rays = cam.width * cam.height
    m = pmarch.Marcher(0)
    fi = m.set_frame(0, scene.xyz, scene.h)
    st_in = scenes.aniso_settings(scene, skip_mode=4)
    wall, kms, st = run_gpu(m, 0, st_in, cam, args.steps, args.warmup)
    B = alg_bytes_aniso(st)
    achieved = B / (kms * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s @1080p/1M particles, plain vs learned-step march, at fixed depth error",
        "value": round(args.steps * rays / wall / 1e6, 3),
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
        "config": {"workload": "anisotropic march (WPCA over 2h, anisotropic kernel; k_n 0.5, k_r 2, k_s 2000, "
                               "N_eps 1) on dam_break 1,024,000 particles scaled to h = 0.1, 1920x1080 view, "
                               "configured march (AABB entry, step h/8, 8 bisections with the last in FP64, "
                               "central-difference normals, exact skip of sub-boxes with no particle within h), "
                               "IsoDensity 1.0",
                   "particles": scene.n, "width": cam.width, "height": cam.height, "h": scene.h,
                   "step": scene.step, "iso": 1.0, "path": "anisotropic"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for("anisotropic", args.workload),
                     "limiter": "VALU issue of the per-lane WPCA and density walks (lane utilisation 42 %); "
                                "achieved/frac are the walks' algorithmic bytes, read mostly from L1/L2"},
        "kernel_ms": round(kms, 4),
        "alg_bytes_per_launch": B,
        "march_counters": {k: st[k] for k in ("rays", "rays_in_aabb", "hits", "samples", "cand", "march_samples")},
        "grid_build_ms": round(float(fi.build_ms), 3),
        "copy_gbps": round(m.stream_copy_gbps(1 << 28, 10), 1),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_aniso(scene, cam, st_in, args.cpu_budget)
        img = m.march(0, st_in, cam.inv_pv, cam.eye, cam.width, cam.height)[:3]
        out["error_vs_fp64"] = fp64_errors(scene, cam, img, 16, settings=st_in)
    m.close()
    print(json.dumps(out))


def bench_reference(args):
    """The reference's own isotropic march -- the facade's path -- on the 1M dam
    break at 1080p: rays start on the GPU sprite pass's depths
    (PerPixel_Isotropic, RayMarcher.cpp:256-344), MaxSteps 128 of the
    reference's step (0.009 at h = 0.1, scaled to h), the DensityGrid box skip,
    normalize(sum gradW), plain FP32 sums."""
    import pmarch
    from pmarch import scenes

    scene = scenes.dam_break(0)
    cam = scene.cameras[0]
    W, H = cam.width, cam.height
    m = pmarch.Marcher(0)
    fi = m.set_frame(0, scene.xyz, scene.h)
    depth = m.depth_sprites(0, cam, scene.h)
    st = dict(mode=1, max_steps=128, step=float(np.float32(0.009 * scene.h / 0.1)), iso=scene.iso,
              normal_mode=1, skip_mode=2, compensated=0)
    for _ in range(args.warmup):
        m.march_async(0, st, cam.inv_pv, cam.eye, W, H, depth=depth, outputs=False)
        m.wait()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.march_async(0, st, cam.inv_pv, cam.eye, W, H, depth=depth, outputs=False)
        m.wait()
        kms.append(m.stats()["kernel_ms"])
    wall = time.perf_counter() - t0
    stats = m.stats()
    kms = float(np.mean(kms))
    B = alg_bytes(stats, "cand") + 4.0 * W * H         # + the depth image
    out = {
        "metric": "Mrays/s @1080p/1M particles, plain vs learned-step march, at fixed depth error",
        "value": round(args.steps * W * H / wall / 1e6, 3),
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
        "config": {"workload": "reference-mode march (PerPixel_Isotropic: start on the sprite depths, MaxSteps "
                               "128, step 0.009 h / 0.1, DensityGrid box skip, normalize(sum gradW), plain FP32 "
                               "sums) of the 1,024,000-particle dam break, 1920x1080 view",
                   "particles": scene.n, "width": W, "height": H, "h": scene.h, "step": st["step"],
                   "iso": scene.iso, "path": "reference"},
        "roofline": {"bound": "valu", "achieved": round(B / (kms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(B / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes": "12 cand (walk candidates) + 72 samples + 32 rays + 4 depth per ray"},
        "kernel_ms": round(kms, 4),
        "march_counters": {k: stats[k] for k in ("rays", "rays_in_aabb", "hits", "samples", "cand")},
        "grid_build_ms": round(float(fi.build_ms), 3),
    }
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O

        threads = max(1, min(16, os.cpu_count() or 2) - 1)
        G = O.Grid(scene.xyz, scene.h)
        ost = O.make_settings(**st)
        step = 1
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            _, _, _, cnt = O.march(G, ost, cam.inv_pv, cam.eye, W, H, depth=depth, row_step=step, threads=threads,
                                   want_t=False)
            runs.append(time.perf_counter() - t0)
            if sum(runs) > args.cpu_budget:
                break
        dt = float(np.median(runs))
        out["cpu_baseline"] = {"value": round(cnt["rays"] / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the same {W}x{H} view ({cnt['rays']} rays) from the same depths, "
                                         f"median of {len(runs)} runs of {dt:.2f} s, {threads} threads on "
                                         f"{os.cpu_count()} cpus ({cpu_model()})"}
    m.close()
    print(json.dumps(out))


def bench_autoplay(args):
    """16 frames of the 1M dam break, each marched once at 1080p while the next
    one streams in, against the serial order."""
    import pmarch
    from pmarch import scenes

    base = scenes.dam_break(0)
    nf = 16
    frames = scenes.displaced_frames(base, nf, seed=1)
    cam = base.cameras[0]
    W, H = cam.width, cam.height
    # a streamed frame is marched once: the bound bits of 4x4x4 sub-boxes per
    # cell, whose build is about 5x cheaper than that of 8x8x8 (same images)
    st = base.settings(skip_mode=4, bound_boxes=1)
    m = pmarch.Marcher(0)
    kms = []

    def march(slot):
        m.march_async(slot, st, cam.inv_pv, cam.eye, W, H, outputs=False)

    def serial():
        for i in range(nf):
            m.set_frame(i % 2, frames[i], base.h)
            march(i % 2)
            m.wait()

    def streamed():
        m.set_frame(0, frames[0], base.h)
        for i in range(nf):
            march(i % 2)
            if i + 1 < nf:
                m.set_frame_stream((i + 1) % 2, frames[i + 1], base.h)
            m.wait()
            kms.append(m.stats()["kernel_ms"])

    def timed(play):
        for _ in range(args.warmup):
            play()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            play()
        return time.perf_counter() - t0

    wall_serial = timed(serial)
    st = base.settings(skip_mode=4, bound_boxes=0)
    wall_fine = timed(streamed)
    st = base.settings(skip_mode=4, bound_boxes=1)
    kms.clear()
    wall = timed(streamed)
    stats = m.stats()
    km = float(np.mean(kms))
    B = alg_bytes(stats, "cand")
    achieved = B / (km * 1e-3) / 1e9
    n = args.steps * nf
    out = {
        "metric": "Mrays/s @1080p/1M particles, a new frame for every march (autoplay)",
        "value": round(n * W * H / wall / 1e6, 3),
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
        "config": {"workload": "autoplay: 16 consecutive frames of the 1,024,000-particle dam break (1 % "
                               "displacement per frame), each uploaded from host memory, built and marched once "
                               "at 1920x1080 (configured march, exact bound skip over 4x4x4 sub-boxes per "
                               "cell); the next frame streams in on the second stream during each march",
                   "particles": base.n, "frames": nf, "width": W, "height": H, "path": "autoplay"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for("plain", "dam_break_1080p")},
        "ms_per_frame": round(wall / n * 1e3, 4),
        "serial_ms_per_frame": round(wall_serial / n * 1e3, 4),
        "overlap_speedup": round(wall_serial / wall, 3),
        "ms_per_frame_8x8x8_boxes": round(wall_fine / n * 1e3, 4),
        "march_kernel_ms": round(km, 4),
        "march_counters": {k: stats[k] for k in ("rays", "rays_in_aabb", "hits", "samples", "cand")},
    }
    m.close()
    print(json.dumps(out))


def sprite_pixel_tests(xyz, cam, R: float) -> int:
    """Pixel centres k_depth_sprites tests in one pass: per particle the pixels
    of its projected quad's bounding box, clamped to the image (particles with
    a corner behind the eye draw nothing) -- the kernel's own loop bounds,
    evaluated in FP32 in the kernel's operation order."""
    f = np.float32
    V = np.asarray(cam.view, f).reshape(16)
    P = np.asarray(cam.proj, f).reshape(16)
    W, H = cam.width, cam.height
    q = np.asarray(xyz, f)
    a = np.array([V[0], V[4], V[8]], f)
    b = np.array([V[1], V[5], V[9]], f)
    lo_x = np.full(len(q), f(3.0e38)); hi_x = np.full(len(q), f(-3.0e38))
    lo_y = lo_x.copy(); hi_y = hi_x.copy()
    ok = np.ones(len(q), bool)
    R = f(R)

    def xform(M, x, y, z):
        return [(M[r] * x + M[4 + r] * y) + (M[8 + r] * z + M[12 + r] * f(1.0)) for r in range(4)]

    for k in range(4):
        sx = f(1.0) if k & 1 else f(-1.0)
        sy = f(1.0) if k & 2 else f(-1.0)
        c = [q[:, i] + R * (sx * a[i] + sy * b[i]) for i in range(3)]
        v = xform(V, *c)
        cl = xform(P, v[0], v[1], v[2])
        ok &= cl[3] > 0
        w3 = np.where(cl[3] > 0, cl[3], f(1.0))
        fx = (cl[0] / w3 + f(1.0)) * f(0.5) * f(W)
        fy = (cl[1] / w3 + f(1.0)) * f(0.5) * f(H)
        lo_x = np.minimum(lo_x, fx); hi_x = np.maximum(hi_x, fx)
        lo_y = np.minimum(lo_y, fy); hi_y = np.maximum(hi_y, fy)
    x0 = np.maximum(np.ceil(lo_x - f(0.5)), f(0.0)); x1 = np.minimum(np.floor(hi_x - f(0.5)), f(W - 1))
    y0 = np.maximum(np.ceil(lo_y - f(0.5)), f(0.0)); y1 = np.minimum(np.floor(hi_y - f(0.5)), f(H - 1))
    n = np.where(ok & (x0 <= x1) & (y0 <= y1), (x1 - x0 + 1).astype(np.int64) * (y1 - y0 + 1).astype(np.int64), 0)
    return int(n.sum())


# FP32 operations per pixel test of k_depth_sprites before its disc test (the
# pixel's ray, its intersection with the quad's plane and the UV disc test);
# a drawn fragment adds about 90 more (two more matrix products, the cosine
# polynomial, the projection) that this count leaves out.
SPRITE_FLOP_PER_TEST = 58
FP32_PEAK_TFLOPS = 157.3


def bench_sprites(args):
    """The depth-sprite pass of the 1M dam break at 1080p into device memory:
    the steps are enqueued back to back on the context stream and timed from
    the first launch to the end of the last pass."""
    import torch
    import pmarch
    from pmarch import scenes

    sc = scenes.dam_break(0)
    cam = sc.cameras[0]
    m = pmarch.Marcher(0)
    m.set_frame(0, sc.xyz, sc.h)
    d = torch.empty(cam.height * cam.width, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        m.depth_sprites_device(0, cam, sc.h, d.data_ptr())
        m.wait()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.depth_sprites_device(0, cam, sc.h, d.data_ptr())
    m.wait()
    wall = time.perf_counter() - t0
    ms = wall / args.steps * 1e3
    tests = sprite_pixel_tests(sc.xyz, cam, sc.h)
    tflops = SPRITE_FLOP_PER_TEST * tests / (ms * 1e-3) / 1e12
    img = d.cpu().numpy()
    out = {
        "metric": "depth sprites Mparticles/s @1080p/1M particles (DepthRenderPass), image in device memory",
        "value": round(args.steps * sc.n / wall / 1e6, 3),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "depth sprites (radius h, camera-facing quads, depth test LESS) of the "
                               "1,024,000-particle dam break, 1920x1080 view, image in device memory",
                   "particles": sc.n, "width": cam.width, "height": cam.height, "path": "sprites"},
        "roofline": {"bound": "valu", "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "traffic": None,
                     "flops": f"{SPRITE_FLOP_PER_TEST} per pixel test x {tests} tests per pass (drawn fragments' "
                              "extra work not counted)"},
        "pixel_tests_per_pass": tests,
        "pixels_drawn": int((img < 1.0).sum()),
        "timing": "host clock over back-to-back passes on the context stream (fill + sprites kernels)",
    }
    m.close()
    print(json.dumps(out))


def smooth_depth_image(W: int, H: int, seed: int = 0) -> np.ndarray:
    """A sprite-style depth image: 1.0 (nothing drawn) outside a disc of smooth
    depth values with per-pixel noise; the pass has no data-dependent work, so
    the content only has to be realistic, not exact."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    r = np.hypot((x - W / 2) / (0.4 * W), (y - H / 2) / (0.4 * H))
    d = 0.6 + 0.2 * r + rng.uniform(-0.01, 0.01, (H, W))
    return np.where(r < 1.0, d, 1.0).astype(np.float32)


def cpu_baseline_smooth(depth: np.ndarray, N: int, budget_s: float):
    """The oracle's smoothing (one thread, the same sums in the same order) on
    the same image."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    H, W = depth.shape
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.depth_smooth(depth, N)
        runs.append(time.perf_counter() - t0)
        if sum(runs) > budget_s:
            break
    dt = float(np.median(runs))
    return {"value": round(W * H / dt / 1e6, 4), "unit": "Mpixels/s", "cores": 1, "kind": "port",
            "sample": f"the same {W}x{H} image, N = {N}, median of {len(runs)} runs of {dt:.2f} s, 1 thread "
                      f"({cpu_model()})"}


def bench_smooth(args):
    """GaussRenderPass's depth smoothing, N = 8, on a 1080p image in HBM."""
    import torch
    import pmarch

    W, H, N = 1920, 1080, 8
    depth = smooth_depth_image(W, H)
    m = pmarch.Marcher(0)
    din = torch.from_numpy(depth).to("cuda:0")
    dout = torch.empty_like(din)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    run = lambda: m.depth_smooth_device(din.data_ptr(), dout.data_ptr(), W, H, N, stream.cuda_stream)
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        run()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    B = 8.0 * W * H                    # one read and one write of the image
    achieved = B / (kms * 1e-3) / 1e9
    tflops = 2.0 * (2 * N + 1) ** 2 * W * H / (kms * 1e-3) / 1e12
    out = {
        "metric": "depth smoothing Mpixels/s @1080p (GaussRenderPass, 17x17 window)",
        "value": round(args.steps * W * H / wall / 1e6, 3),
        "unit": "Mpixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "screen-space depth smoothing (GaussRenderPass: normalised Gaussian, N = 8, "
                               "clamp-to-edge) of a 1920x1080 depth image in device memory",
                   "width": W, "height": H, "N": N, "path": "smooth"},
        "roofline": {"bound": "valu", "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "traffic": traffic_for("smooth", args.workload),
                     "flops": f"2 x {(2 * N + 1) ** 2} per pixel (one fused multiply-add per window tap)",
                     "hbm_gbps": round(achieved, 2), "hbm_frac": round(achieved / HBM_PEAK_GBS, 4)},
        "kernel_ms": round(kms, 5),
        "alg_bytes_per_launch": B,
        "copy_gbps": round(m.stream_copy_gbps(1 << 28, 10), 1),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_smooth(depth, N, args.cpu_budget)
    m.close()
    print(json.dumps(out))


def bench_batch(args):
    """64 views x 16 frames in one persistent launch from a hipGraph."""
    import pmarch
    from pmarch import scenes, train

    base = scenes.dam_break(0)
    frames = scenes.displaced_frames(base, 16, seed=1)
    rng = np.random.default_rng(2024)
    cams = scenes.random_views(rng, scenes.COLUMN_CENTRE, 64, 256, 1920, 256, 1080, radius=(1.1, 1.9))
    m = pmarch.Marcher(0)
    build_ms = []
    for f, xyz in enumerate(frames):
        build_ms.append(float(m.set_frame(f, xyz, base.h).build_ms))
    views = [(f, c.width, c.height, c.inv_pv, c.eye) for c in cams for f in range(16)]
    rays = sum(W * H for _, W, H, _, _ in views)
    forms = {}
    plain = base.settings(skip_mode=4)
    sets = {"plain": plain}
    if not args.plain_only and train.WEIGHTS.exists():
        m.mlp_set(train.load()[0])
        sets["learned"] = base.settings(learned=1, skip_mode=4)
    for name, st in sets.items():
        m.batch_set(st, views, use_graph=True)
        for _ in range(args.warmup):
            m.batch_run()
            m.wait()
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            m.batch_run()
            m.wait()
            kms.append(m.stats()["kernel_ms"])
        wall = time.perf_counter() - t0
        forms[name] = dict(mrays=args.steps * rays / wall / 1e6, ms=wall / args.steps * 1e3,
                           kernel_ms=float(np.mean(kms)), stats=m.stats())
    # every jump of the learned form is certified: its images equal the plain ones
    learned_par = None
    if "learned" in forms:
        check = list(range(0, len(views), 61))
        m.batch_set(plain, views, use_graph=False)
        m.batch_run()
        m.wait()
        ref = {i: m.batch_download(i) for i in check}
        m.batch_set(sets["learned"], views, use_graph=False)
        m.batch_run()
        m.wait()
        same = [all(np.array_equal(a, b) for a, b in zip(m.batch_download(i), ref[i])) for i in check]
        learned_par = {"views_checked": len(check), "identical_images": bool(all(same))}
    use = max(forms, key=lambda k: forms[k]["mrays"])
    f = forms[use]
    st = f["stats"]
    # the batch kernel has no instrumented form: its bytes use the candidates its
    # walks read (the six central differences' shared walk counted once)
    B = alg_bytes(st, "cand")
    achieved = B / (f["kernel_ms"] * 1e-3) / 1e9
    out = {
        "metric": "Mrays/s @1080p/1M particles, plain vs learned-step march, at fixed depth error",
        "value": round(f["mrays"], 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(f["ms"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch: 64 views (256..1920 x 256..1080) x 16 frames of the 1,024,000-particle dam "
                               "break, one persistent launch replayed from a hipGraph", "views": 64, "frames": 16,
                   "rays_per_step": rays, "path": use},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(use, "batch"),
                     "bytes": "12 cand (walk candidates) + 72 samples + 32 rays"},
        "forms": {k: {"mrays_per_s": round(v["mrays"], 3), "kernel_ms": round(v["kernel_ms"], 3),
                      "samples": v["stats"]["samples"], "march_samples": v["stats"]["march_samples"],
                      "mlp_evals": v["stats"]["mlp_evals"], "cert_skipped": v["stats"]["cert_skipped"]}
                  for k, v in forms.items()},
        "learned_vs_plain": learned_par,
        "alg_bytes_per_launch": B,
        "frame_build_ms_mean": round(float(np.mean(build_ms)), 3),
    }
    if not args.no_cpu:
        c0 = cams[0]
        out["cpu_baseline"] = cpu_baseline(base, c0, base.settings(), args.cpu_budget)
        # the oracle leg's check: four view-frames spread over cameras and
        # frames against the FP64 march of their frame, every 32nd row
        errs = []
        for i in (0, 341, 682, 1023):
            errs.append(fp64_errors(base, cams[i // 16], m.batch_download(i), 32, xyz=frames[i % 16]))
        out["error_vs_fp64"] = merge_errors(errs)
    m.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["dam_break_1080p", "clustered_4k", "batch", "aniso_dam_1080p",
                                           "depth_smooth_1080p", "autoplay_dam_1080p", "sprites_dam_1080p",
                                           "reference_dam_1080p"],
                    default="dam_break_1080p")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "aniso_dam_1080p":
        bench_aniso(args)
        return
    if args.workload == "depth_smooth_1080p":
        bench_smooth(args)
        return
    if args.workload == "autoplay_dam_1080p":
        bench_autoplay(args)
        return
    if args.workload == "sprites_dam_1080p":
        bench_sprites(args)
        return
    if args.workload == "reference_dam_1080p":
        bench_reference(args)
        return

    import pmarch
    from pmarch import scenes, train

    if args.workload == "clustered_4k":
        scene = scenes.clustered_pool(0)
        wdesc = ("clustered pool 4,000,000 particles (pool layer, 200 droplets, splash cones), 3840x2160 view, "
                 "configured march (AABB entry, step r/2, 8 bisections with the last in FP64, central-difference "
                 "normals, compensated FP32 sums)")
        fp64_step = 72
    else:
        scene = scenes.dam_break(0)
        wdesc = ("dam_break 1,024,000 particles, 1920x1080 view, configured march (AABB entry, step r/2, 8 "
                 "bisections with the last in FP64, central-difference normals, compensated FP32 sums)")
        fp64_step = 27
    cam = scene.cameras[0]
    rays = cam.width * cam.height
    m = pmarch.Marcher(0)
    fi = m.set_frame(0, scene.xyz, scene.h)

    # The forms, all with the same image (the skips and the certified jumps
    # pass only samples proven below iso): the configured march as written
    # (fixed step from the AABB entry), with the exact sub-box bound skip, and
    # the learned step proposal on top of the bound skip.
    sets = {"fixed_step": scene.settings(skip_mode=0), "plain": scene.settings(skip_mode=4)}
    if not args.plain_only and train.WEIGHTS.exists():
        m.mlp_set(train.load()[0])
        sets["learned"] = scene.settings(learned=1, skip_mode=4)
    forms, images = {}, {}
    for name, st in sets.items():
        wall, kms, stats = run_gpu(m, 0, st, cam, args.steps, args.warmup)
        ins = instrumented(m, 0, st, cam)
        forms[name] = dict(mrays=args.steps * rays / wall / 1e6, ms=wall / args.steps * 1e3, kernel_ms=kms,
                           stats=stats, ins=ins)
        images[name] = m.march(0, st, cam.inv_pv, cam.eye, cam.width, cam.height)[:3]
    same = {k: bool(all(np.array_equal(a, b) for a, b in zip(images[k], images["fixed_step"]))) for k in images}
    fi = m.frame_info(0)
    # the PCIe-inclusive rate: the same march with the images copied to host memory
    m.march(0, sets["plain"], cam.inv_pv, cam.eye, cam.width, cam.height)
    pcie_ms = m.stats()["total_ms"]

    use = max(forms, key=lambda k: forms[k]["mrays"])
    f = forms[use]
    B = alg_bytes(f["ins"])
    B27 = alg_bytes(f["ins"], "cand27")
    kt = f["kernel_ms"] * 1e-3
    copy = m.stream_copy_gbps(1 << 28, 10)
    traffic = traffic_for(use if use != "fixed_step" else "plain", args.workload)
    out = {
        "metric": "Mrays/s @1080p/1M particles, plain vs learned-step march, at fixed depth error",
        "value": round(f["mrays"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wdesc, "particles": scene.n, "width": cam.width, "height": cam.height, "h": scene.h,
                   "step": scene.step, "iso": scene.iso, "path": use},
        "roofline": {"bound": "valu", "achieved": round(B / kt / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(B / kt / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes": "12 cand_stencil + 72 samples + 32 rays (instrumented launch of the same view)",
                     "frac_27cell": round(B27 / kt / 1e9 / HBM_PEAK_GBS, 4),
                     "frac_of_copy": round(B / kt / 1e9 / copy, 4),
                     "hbm_gbps": round(traffic / kt / 1e9, 2) if traffic else None,
                     "hbm_frac": round(traffic / kt / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "limiter": "VALU issue of divergent per-lane stencil walks (SIMD VALU busy ~79 %, lane "
                                "utilisation 47 %); achieved/frac are the algorithmic bytes of the HBM "
                                "roofline, read mostly from L1/L2 -- the HBM traffic itself is hbm_gbps"},
        "kernel_ms": round(f["kernel_ms"], 4),
        "alg_bytes_per_launch": B,
        "alg_bytes_27cell_per_launch": B27,
        "forms": {k: {"mrays_per_s": round(v["mrays"], 3), "kernel_ms": round(v["kernel_ms"], 4),
                      "samples": v["stats"]["samples"], "march_samples": v["stats"]["march_samples"],
                      "mlp_evals": v["stats"]["mlp_evals"], "cert_walks": v["stats"]["cert_walks"],
                      "cert_skipped": v["stats"]["cert_skipped"],
                      "frac": round(alg_bytes(v["ins"]) / (v["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "image_equals_fixed_step": same[k]} for k, v in forms.items()},
        "learned_vs_plain": (round(forms["learned"]["mrays"] / forms["plain"]["mrays"], 3) if "learned" in forms
                             else None),
        "learned_vs_fixed_step": (round(forms["learned"]["mrays"] / forms["fixed_step"]["mrays"], 3)
                                  if "learned" in forms else None),
        "march_counters": {k: f["ins"][k] for k in ("rays", "rays_in_aabb", "hits", "samples", "march_samples",
                                                     "cand", "cand_stencil", "cand27")},
        "pcie_inclusive_mrays": round(rays / pcie_ms / 1e3, 3),
        "grid_build_ms": round(float(fi.build_ms), 3),
        "bound_build_ms": round(float(fi.bound_ms), 3),
        "copy_gbps": round(copy, 1),
    }
    if not args.no_cpu:
        # the oracle leg: the reference's algorithm on the CPU (configured march
        # without the bound skip), and the FP64 march that checks the GPU image
        out["cpu_baseline"] = cpu_baseline(scene, cam, scene.settings(), args.cpu_budget)
        out["error_vs_fp64"] = dict(fp64_errors(scene, cam, images[use], fp64_step), forms_identical=all(same.values()))
    m.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
