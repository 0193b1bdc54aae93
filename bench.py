"""Benchmark: Mrays/s of the 1024^2 Kerr Stokes render on one MI355X.

Workload: a=0.94, i=17 deg, 1024x1024 pixels over +-20 M from r=1000, adaptive
RK2 h = 0.05 (r - r_+) capped at 4096 steps, 128x64x128 11-channel brick with
a thin disk.  A "step" of this benchmark is one full render of that image with
the brick already resident in HBM.  Each path's error is measured against the
full-FP64 GPU render of the same scene (itself checked against the CPU oracle
by the tests).  The CPU baseline is the NumPy FP64 oracle on a bounded,
evenly spaced row sample, run on all host cores.

Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "blackhole-sim_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

import numpy as np  # noqa: E402

# Algorithmic FP32 work of the fused render, counted from bh_render.hip
# (FMA = 2 FLOP; sin, cos, rcp, sqrt and log2 count 1 each):
#   one RK2 geodesic step (step size, two Carter-form right-hand sides, midpoint
#   and update, transfer midpoint): F_GEO; one in-brick transfer segment
#   (three brackets, trilinear weights, 11-channel interpolation, ds, RK2 Stokes
#   step, optical depth): F_SAMPLE.  The kernel counts steps and samples exactly.
F_GEO = 155.0
F_SAMPLE = 298.0
PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 (packed); scalar issue caps at half of it
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the plain FP32 kernel from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled x2.0 / x1.0 by a 256 MiB copy); profiles/render_kerr_stokes_1024.md
TRAFFIC_PLAIN_FP32 = 1.836e9 + 27.4e6


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def rel_l1(a, b) -> float:
    return float(np.sum(np.abs(a - b)) / max(float(np.sum(np.abs(b))), 1e-300))


def pixel_rel_I(I, I64, ok) -> dict:
    """Per-pixel relative error on I over unflagged pixels whose FP64 I is not negligible
    (|I64| > 1e-6 of the image's largest; pixels whose rays never sample the brick have I = 0)."""
    a, b = np.asarray(I, np.float64)[ok], np.asarray(I64, np.float64)[ok]
    keep = np.abs(b) > 1e-6 * max(float(np.abs(b).max()), 1e-300)
    rel = np.abs(a[keep] - b[keep]) / np.abs(b[keep])
    if rel.size == 0:
        return {"pixels": 0}
    return {"pixels": int(rel.size), "median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
            "max": float(rel.max())}


def cpu_baseline(scene, brick, seconds: float) -> dict:
    """Time the FP64 NumPy oracle on evenly spaced image rows using every host core."""
    from oracle import render as orender

    cores = len(os.sched_getaffinity(0))
    W = scene.width
    t = time.perf_counter()
    calib_rows = np.array([scene.height // 2])
    orender.render_rays(scene, brick, (calib_rows[:, None] * W + np.arange(W)).ravel())
    per_row = time.perf_counter() - t
    n_rows = int(max(cores, min(scene.height, seconds * cores / max(per_row, 1e-6))))
    n_rows = max(cores, min(n_rows, scene.height))
    rows = np.unique(np.linspace(0, scene.height - 1, n_rows).round().astype(np.int64))
    pix = (rows[:, None] * W + np.arange(W)).ravel()
    t = time.perf_counter()
    orender.render_rays_parallel(scene, brick, pix, processes=cores)
    dt = time.perf_counter() - t
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": pix.size / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"FP64 NumPy oracle, {rows.size} evenly spaced rows of {scene.height} "
                      f"({pix.size} rays), {dt:.1f} s on {cores} processes; CPU: {cpu}"}


def time_path(rr, scene, db, precision, bufs, steps: int, warmup: int) -> dict:
    """Wall time per render over `steps` renders enqueued back to back (one synchronisation at the end),
    then the kernel's own time (HIP events) and work counters over as many renders, read after each.
    The warm-up renders also run back to back: the GPU's clocks take ~20 renders of 1024^2 (~40 ms of
    work) to settle, and the kernel time falls by ~9 % over them."""
    d_iquv, d_steps, d_status = bufs
    for _ in range(warmup):
        rr.render_device(scene, db, precision, d_iquv, d_steps, d_status)
    rr.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rr.render_device(scene, db, precision, d_iquv, d_steps, d_status)
    rr.synchronize()
    t1 = time.perf_counter()
    kms = []
    work = []
    for _ in range(steps):
        rr.render_device(scene, db, precision, d_iquv, d_steps, d_status)
        kms.append(rr.last_kernel_ms())
        work.append(rr.last_work())
    rr.synchronize()
    return {"ms_per_step": (t1 - t0) * 1e3 / steps, "kernel_ms": float(np.mean(kms)), "work": work[-1]}


def sampler_path(rr, brick, n: int, reps: int) -> dict:
    """FP64 trilinear sampler (sample_brick_trilinear, device-resident) on n uniform points inside the
    brick: HBM-bound; algorithmic bytes per point = 24 (coordinates) + 88 (11 FP64 channels out)."""
    import ctypes as C
    from bhgrrt import _lib
    from oracle import micro

    rng = np.random.default_rng(20240607)
    pts = np.empty((n, 3))
    pts[:, 0] = rng.uniform(float(brick.r[0]), float(brick.r[-1]), n)
    pts[:, 1] = rng.uniform(float(brick.theta[0]), float(brick.theta[-1]), n)
    pts[:, 2] = rng.uniform(0.0, 2.0 * math.pi, n)
    coeffs = np.ascontiguousarray(brick.coeffs, dtype=np.float64)
    grids = [np.ascontiguousarray(g, dtype=np.float64) for g in (brick.r, brick.theta, brick.phi)]
    bufs = [rr.alloc(a.nbytes) for a in [coeffs, *grids, pts]]
    for b, a in zip(bufs, [coeffs, *grids, pts]):
        b.from_host(a)
    d_out = rr.alloc(n * 11 * 8)
    lib = rr._lib

    def run():
        _lib.check(lib.bh_sample_brick_trilinear_device(rr.ctx, bufs[0].ptr, bufs[1].ptr, grids[0].size, bufs[2].ptr,
                                                        grids[1].size, bufs[3].ptr, grids[2].size, bufs[4].ptr,
                                                        C.c_int64(n), d_out.ptr))

    run()
    rr.synchronize()
    kms = []
    for _ in range(max(1, reps)):
        run()
        kms.append(rr.last_kernel_ms())
    rr.synchronize()
    out = d_out.to_host(np.empty((n, 11)))
    pick = rng.choice(n, 4096, replace=False)
    ref, _ = micro.sample_vec(coeffs, *grids, pts[pick, 0], pts[pick, 1], pts[pick, 2])
    err = float(np.max(np.abs(out[pick] - ref) / np.maximum(np.abs(ref), 1e-300)))
    # the reference's parity gate (native_kernels.py:12-15): |a - b| <= ATOL + RTOL |b|, RTOL = ATOL = 1e-12
    gate = float(np.max(np.abs(out[pick] - ref) / (1e-12 + 1e-12 * np.abs(ref))))
    ms = float(np.mean(kms))
    gbs = n * 112.0 / (ms * 1e-3) / 1e9
    for b in bufs + [d_out]:
        b.free()
    return {"points": n, "kernel_ms": ms, "gpoints_s": n / (ms * 1e-3) / 1e9, "algorithmic_GBs": gbs,
            "hbm_frac": gbs / PEAK_HBM_GBS, "bytes_per_point": 112, "max_rel_err_vs_oracle": err,
            "checked_points": int(pick.size), "allclose_1e-12": gate <= 1.0, "allclose_margin": gate}


def micro_path(calls: int) -> dict:
    """The small configuration's sampler micro-benchmark: sample_and_step_stokes (the drop-in for
    blackhole_native, invalid_policy='zero') on the 8x6x8 brick at 64 seeded uniform points (r and theta
    drawn 10 % beyond the brick so the mask and the policy run), one iteration per call as the reference's
    benchmark defaults (benchmark_cli.py:18-22).  Host arrays in and out: the figure is the call's latency
    (validation, copies, launch, synchronisation), reported as the median of ``calls`` calls; the FP64
    oracle's NumPy time for the same call is the CPU figure."""
    from bhgrrt import native_kernels as nk
    from bhgrrt import scenes
    from oracle import micro

    br = scenes.small_brick()
    coeffs = br.coeffs.astype(np.float64)
    rg, tg, pg = (g.astype(np.float64) for g in (br.r, br.theta, br.phi))
    rng = np.random.default_rng(20240605)
    n = 64
    pts = np.empty((n, 3))
    span_r, span_t = rg[-1] - rg[0], tg[-1] - tg[0]
    pts[:, 0] = rng.uniform(rg[0] - 0.1 * span_r, rg[-1] + 0.1 * span_r, n)
    pts[:, 1] = rng.uniform(tg[0] - 0.1 * span_t, tg[-1] + 0.1 * span_t, n)
    pts[:, 2] = rng.uniform(0.0, 2.0 * math.pi, n)
    init = rng.normal(size=(n, 4))
    ds = 0.37

    def gpu():
        return nk.sample_and_step_stokes(coeffs, rg, tg, pg, pts, ds, init, invalid_policy="zero")

    def cpu():
        return micro.sample_and_step_stokes_vec(coeffs, rg, tg, pg, pts, ds, init, invalid_policy="zero")

    out, ref = gpu(), cpu()
    gate = float(np.max(np.abs(out - ref) / (1e-12 + 1e-12 * np.abs(ref))))
    t_gpu, t_cpu = [], []
    for _ in range(calls):
        t0 = time.perf_counter()
        gpu()
        t_gpu.append(time.perf_counter() - t0)
    for _ in range(max(1, calls // 10)):
        t0 = time.perf_counter()
        cpu()
        t_cpu.append(time.perf_counter() - t0)
    valid = nk.sample_brick_valid_mask(rg, tg, pts)
    return {"points": n, "iterations_per_call": 1, "calls": calls, "latency_us_median": float(np.median(t_gpu)) * 1e6,
            "latency_us_p90": float(np.quantile(t_gpu, 0.9)) * 1e6, "invalid_points": int(n - valid.sum()),
            "allclose_1e-12_vs_oracle": gate <= 1.0, "allclose_margin": gate,
            "cpu_oracle_latency_us_median": float(np.median(t_cpu)) * 1e6}


def precompute_path(rr, dims, reps: int, cpu_cells: int) -> dict:
    """Coefficient-brick precompute (precompute_coefficient_bricks) on the analytic torus: device-resident
    fields, float32 brick out; 1024 sampled cells checked against the FP64 oracle from a float64 launch,
    and the oracle's own rate on ``cpu_cells`` cells as the CPU figure."""
    import ctypes as C
    from bhgrrt import _lib
    from bhgrrt import coefficient_bricks as cb
    from bhgrrt.grmhd import generate_analytic_grmhd_torus
    from oracle import synchro as osy

    snap = generate_analytic_grmhd_torus(spin_a=0.85, nr=dims[0], ntheta=dims[1], nphi=dims[2])
    sc = cb.PhysicalScaling.from_mdot(cb.BlackHoleSystem.m87_star(0.85), 1e-4)
    n = snap.rho.size
    fields = [np.ascontiguousarray(a) for a in (snap.rho, snap.theta_e, snap.b_con)]
    bufs = [rr.alloc(a.nbytes) for a in fields]
    for b, a in zip(bufs, fields):
        b.from_host(a)
    d_out = rr.alloc(n * 11 * 8)
    desc = cb.synchrotron_desc(sc, 230e9)

    def run(dtype):
        _lib.check(rr._lib.bh_precompute_coefficient_bricks_device(rr.ctx, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                                                   C.c_int64(n), C.byref(desc), _lib.BRICK_DTYPE[dtype],
                                                                   d_out.ptr))

    run("float64")
    rr.synchronize()
    out64 = d_out.to_host(np.empty((n, 11)))
    run("float32")
    rr.synchronize()
    kms = []
    for _ in range(max(1, reps)):
        run("float32")
        kms.append(rr.last_kernel_ms())
    rr.synchronize()
    for b in bufs + [d_out]:
        b.free()
    rng = np.random.default_rng(11)
    pick = rng.choice(n, 1024, replace=False)
    take = lambda a: a.reshape(n, *a.shape[3:])[pick].reshape(1024, 1, 1, *a.shape[3:])
    ref = osy.precompute_coefficient_bricks(take(snap.rho), take(snap.theta_e), take(snap.b_con), sc.rho_cgs_per_code,
                                            sc.b_gauss_per_code, 230e9, "float64").reshape(1024, 11)
    nz = ref != 0
    err = float(np.max(np.abs(out64[pick][nz] - ref[nz]) / np.abs(ref[nz])))
    cpu = {}
    if cpu_cells > 0:
        sub = rng.choice(n, cpu_cells, replace=False)
        takes = lambda a: a.reshape(n, *a.shape[3:])[sub].reshape(cpu_cells, 1, 1, *a.shape[3:])
        t0 = time.perf_counter()
        osy.precompute_coefficient_bricks(takes(snap.rho), takes(snap.theta_e), takes(snap.b_con), sc.rho_cgs_per_code,
                                          sc.b_gauss_per_code, 230e9, "float32")
        dt = time.perf_counter() - t0
        cpu = {"value": cpu_cells / dt / 1e6, "unit": "Mcells/s", "cores": 1, "kind": "port",
               "sample": f"{cpu_cells} seeded cells of the same torus through the NumPy oracle"}
    ms = float(np.mean(kms))
    # per cell: 8 distinct pitch nodes x (192 thermal + 224 power-law Lorentz factors) table evaluations
    evals = 8 * (192 + 224)
    return {"grid": list(dims), "cells": n, "kernel_ms": ms, "mcells_s": n / (ms * 1e-3) / 1e6,
            "gevals_s": n * evals / (ms * 1e-3) / 1e9, "evaluations_per_cell": evals, "dtype_out": "float32",
            "max_rel_err_vs_oracle": err, "cpu_oracle": cpu}


def surrogate_path(rr, scene, brick, ref, I64, args, frame_scenes, frames64) -> tuple[dict, dict]:
    """Train the MLP surrogate on seeded RK2 geodesics, then time surrogate renders of the scene
    (bf16 corner brick, mixed-precision RK2 fallback) and measure their error against FP64: one render
    at a time, and the batched frames in one captured sequence (inputs, one FP16 inference, the knot
    transfer of every frame, all frames' fallback rays in one persistent RK2 launch)."""
    from bhgrrt import surrogate as bsur
    from bhgrrt.render import PREC_MIXED

    cfg = bsur.SurrogateConfig(fallback_precision=PREC_MIXED)
    log(f"[bench] surrogate: {args.surrogate_rays} RK2 geodesics, Adam lr 1e-3, batch 8192, "
        f"<= {args.surrogate_epochs} epochs / {args.surrogate_seconds} s")
    t0 = time.perf_counter()
    sur, info = bsur.train_surrogate(rr, brick, cfg, n_rays=args.surrogate_rays, epochs=args.surrogate_epochs,
                                     time_limit_s=args.surrogate_seconds)
    train_s = time.perf_counter() - t0
    log(f"[bench] surrogate trained: {info['steps']} steps in {info['seconds']:.1f} s, loss {info['loss']:.3f}, "
        f"flagged {info['flagged']:.3f}")
    dbx = rr.upload(brick, "bf16", "corner")
    n = scene.rays
    bufs = (rr.alloc(n * 16), rr.alloc(n * 4), rr.alloc(n))
    for _ in range(max(1, args.warmup)):
        sur.render_device(scene, dbx, *bufs)
    rr.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sur.render_device(scene, dbx, *bufs)
        kms.append(rr.last_kernel_ms())
    rr.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    fallback = sur.last_fallback()
    iquv = bufs[0].to_host(np.empty((scene.height, scene.width, 4), np.float32))
    status = bufs[2].to_host(np.empty((scene.height, scene.width), np.uint8))
    ok = ((ref["status"] | status) & 8) == 0
    out = {"mrays_s": n / dt / 1e6, "ms_per_render": dt * 1e3, "kernel_ms": float(np.mean(kms)),
           "rel_l1_I": rel_l1(iquv[..., 0][ok].astype(np.float64), I64[ok]),
           "max_abs_QUV": float(np.max(np.abs(iquv[..., 1:][ok].astype(np.float64) - ref["iquv"][..., 1:][ok]))),
           "pixel_rel_I": pixel_rel_I(iquv[..., 0], I64, ok), "fallback_frac": fallback / n, "train_steps": info["steps"], "train_s": round(train_s, 2),
           "train_loss": info["loss"], "train_rays": info["n_rays"], "flagged_train_frac": info["flagged"],
           "knots": cfg.n_knots, "width": cfg.width, "depth": cfg.depth}
    log(f"[bench] surrogate_mixed_bf16: {out}")
    for b in bufs:
        b.free()
    # batched: the frames of the batched paths
    sb = sur.batch([(fs, dbx) for fs in frame_scenes])
    sb.launch(graph=True)
    rr.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        sb.launch(graph=True)
    rr.synchronize()
    dtb = (time.perf_counter() - t0) / max(1, args.steps // 2)
    kmb = rr.last_kernel_ms()
    fb_steps, fb_samples = rr.last_work()
    outs = sb.results()
    per, tot = sb.fallback()
    num = den = 0.0
    pix = []
    for o, r in zip(outs, frames64):
        okb = ((o["status"] | r["status"]) & 8) == 0
        Ib, I0 = o["iquv"][..., 0].astype(np.float64), r["iquv"][..., 0]
        num += float(np.abs(Ib - I0)[okb].sum())
        den += float(np.abs(I0[okb]).sum())
        big = okb & (np.abs(I0) > 1e-6 * np.abs(I0[okb]).max())
        pix.append(np.abs(Ib - I0)[big] / np.abs(I0[big]))
    pix = np.concatenate(pix)
    nb = sb.pixels
    batched = {"mrays_s": nb / dtb / 1e6, "ms_per_launch": dtb * 1e3, "kernel_ms": kmb, "frames": len(frame_scenes),
               "rel_l1_I": num / den, "pixel_rel_I": {"median": float(np.median(pix)),
                                                       "p99": float(np.quantile(pix, 0.99))},
               "fallback_frac": tot / nb, "knots": "fp16", "graph": True,
               "fallback_rk2_steps": fb_steps, "fallback_brick_samples": fb_samples}
    log(f"[bench] surrogate_batched: {batched}")
    sb.close()
    rr.release(dbx)
    sur.close()
    return out, batched


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--frames", type=int, default=8, help="1024^2 frames per batched launch")
    ap.add_argument("--surrogate-rays", type=int, default=1 << 20, help="RK2 geodesics the surrogate trains on")
    ap.add_argument("--surrogate-epochs", type=float, default=1e6, help="epoch cap (the time limit binds first)")
    ap.add_argument("--surrogate-seconds", type=float, default=90.0, help="training time limit")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sampler-points", type=int, default=1 << 24)
    ap.add_argument("--precompute-grid", default="128x64x128", help="analytic torus nr x ntheta x nphi for the "
                    "coefficient precompute")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    from bhgrrt import scenes
    from bhgrrt.render import PREC_FP32, PREC_FP64, PREC_MIXED, Renderer

    scene = scenes.scale_plain_scene(args.size, args.size)
    brick = scenes.scale_brick()
    n = scene.rays
    log(f"[bench] scene {scene.width}x{scene.height}, brick {brick.shape}")
    rr = Renderer(0)
    db32 = rr.upload(brick, "f32")
    d_iquv = rr.alloc(n * 4 * 8)
    d_steps = rr.alloc(n * 4)
    d_status = rr.alloc(n)
    bufs = (d_iquv, d_steps, d_status)

    log("[bench] FP64 reference render")
    ref = rr.render(scene, db32, PREC_FP64)
    I64 = ref["iquv"][..., 0]
    fp64_ms = rr.last_kernel_ms()

    log("[bench] plain FP32 path (corner-packed f32 brick)")
    dbc = rr.upload(brick, "f32", "corner")
    plain = time_path(rr, scene, dbc, PREC_FP32, bufs, args.steps, args.warmup)
    out32 = rr.render(scene, dbc, PREC_FP32)
    # rays flagged (status bit 3) in either render left the geodesic at a pole: excluded, counted
    ok_plain = ok = ((ref["status"] | out32["status"]) & 8) == 0
    err32 = rel_l1(out32["iquv"][..., 0][ok].astype(np.float64), I64[ok])
    quv32 = float(np.max(np.abs(out32["iquv"][..., 1:][ok].astype(np.float64) - ref["iquv"][..., 1:][ok])))
    steps_total, samples_total = plain["work"]
    flops = steps_total * F_GEO + samples_total * F_SAMPLE
    achieved = flops / (plain["kernel_ms"] * 1e-3) / 1e12
    value = n / (plain["ms_per_step"] * 1e-3) / 1e6
    log(f"[bench] plain fp32: {plain['ms_per_step']:.3f} ms/render, kernel {plain['kernel_ms']:.3f} ms, "
        f"{value:.1f} Mrays/s, rel L1 I {err32:.2e}")

    paths = {"plain_fp32": {"mrays_s": value, "ms_per_render": plain["ms_per_step"], "kernel_ms": plain["kernel_ms"],
                            "rel_l1_I": err32, "max_abs_QUV": quv32,
                            "pixel_rel_I": pixel_rel_I(out32["iquv"][..., 0], I64, ok), "tflops": achieved,
                            "layout": "corner"},
             "fp64_reference": {"kernel_ms": fp64_ms, "mrays_s": n / (fp64_ms * 1e-3) / 1e6}}
    log("[bench] plain FP32 path (SoA f32 brick)")
    soa = time_path(rr, scene, db32, PREC_FP32, bufs, args.steps, args.warmup)
    paths["plain_fp32_soa"] = {"mrays_s": n / (soa["ms_per_step"] * 1e-3) / 1e6, "ms_per_render": soa["ms_per_step"],
                               "kernel_ms": soa["kernel_ms"], "layout": "soa"}
    # the small plain scene (576 rays, 64 fixed steps): a latency figure, one launch per render
    log("[bench] small plain scene (32x18, 64 fixed RK2 steps, 8x6x8 brick)")
    s_scene, s_brick = scenes.small_plain_scene(), scenes.small_brick()
    dbs32, dbs64 = rr.upload(s_brick, "f32", "corner"), rr.upload(s_brick, "f32")
    sm32 = time_path(rr, s_scene, dbs32, PREC_FP32, bufs, args.steps, args.warmup)
    sm64 = time_path(rr, s_scene, dbs64, PREC_FP64, bufs, args.steps, args.warmup)
    paths["small_plain"] = {"rays": s_scene.rays, "fp32_kernel_ms": sm32["kernel_ms"],
                            "fp32_ms_per_render": sm32["ms_per_step"], "fp64_kernel_ms": sm64["kernel_ms"],
                            "fp64_ms_per_render": sm64["ms_per_step"]}
    log(f"[bench] small plain: {paths['small_plain']}")
    rr.release(dbs32)
    rr.release(dbs64)
    for storage in ("bf16", "f16"):
        log(f"[bench] mixed precision path, {storage} brick")
        dbx = rr.upload(brick, storage, "corner")
        mx = time_path(rr, scene, dbx, PREC_MIXED, bufs, args.steps, args.warmup)
        outx = rr.render(scene, dbx, PREC_MIXED)
        okx = ((ref["status"] | outx["status"]) & 8) == 0
        st, sm = mx["work"]
        paths[f"mixed_{storage}"] = {
            "mrays_s": n / (mx["ms_per_step"] * 1e-3) / 1e6, "ms_per_render": mx["ms_per_step"],
            "kernel_ms": mx["kernel_ms"],
            "rel_l1_I": rel_l1(outx["iquv"][..., 0][okx].astype(np.float64), I64[okx]),
            "max_abs_QUV": float(np.max(np.abs(outx["iquv"][..., 1:][okx].astype(np.float64) - ref["iquv"][..., 1:][okx]))),
            "pixel_rel_I": pixel_rel_I(outx["iquv"][..., 0], I64, okx),
            "pole_flagged_frac": float(1 - okx.mean()), "rk2_steps": st, "brick_samples": sm}
        log(f"[bench] mixed {storage}: {paths[f'mixed_{storage}']}")
        rr.release(dbx)

    # ---- exact Stokes step (use_exact_matrix_step) on the same scene ------------------------
    log("[bench] exact Stokes step: FP32 corner render against the FP64 exact render")
    scene_x = dataclasses.replace(scene, exact_step=True)
    ref_x = rr.render(scene_x, db32, PREC_FP64)
    fp64x_ms = rr.last_kernel_ms()
    ex = time_path(rr, scene_x, dbc, PREC_FP32, bufs, args.steps, args.warmup)
    out_x = rr.render(scene_x, dbc, PREC_FP32)
    okx = ((ref_x["status"] | out_x["status"]) & 8) == 0
    paths["plain_fp32_exact_step"] = {
        "mrays_s": n / (ex["ms_per_step"] * 1e-3) / 1e6, "ms_per_render": ex["ms_per_step"], "kernel_ms": ex["kernel_ms"],
        "rel_l1_I": rel_l1(out_x["iquv"][..., 0][okx].astype(np.float64), ref_x["iquv"][..., 0][okx]),
        "pixel_rel_I": pixel_rel_I(out_x["iquv"][..., 0], ref_x["iquv"][..., 0], okx),
        "fp64_exact_kernel_ms": fp64x_ms,
        "rk2_vs_exact_rel_l1_I": rel_l1(ref["iquv"][..., 0][okx], ref_x["iquv"][..., 0][okx])}
    log(f"[bench] exact step: {paths['plain_fp32_exact_step']}")

    # ---- batched paths -----------------------------------------------------------------
    from bhgrrt.autotune import autotune
    from bhgrrt.render import Batch

    frames = args.frames
    log(f"[bench] batched: {frames} frames of the {scene.width}^2 scene (camera azimuth steps), one launch")
    frame_scenes = [dataclasses.replace(scene, camera=dataclasses.replace(scene.camera, phi=2.0 * math.pi * k / frames))
                    for k in range(frames)]
    for name, storage, prec in (("batched_fp32", "f32", PREC_FP32), ("batched_mixed_bf16", "bf16", PREC_MIXED),
                                ("batched_mixed_f16", "f16", PREC_MIXED)):
        dbx = rr.upload(brick, storage, "corner")
        bt = Batch(rr, [(fs, dbx) for fs in frame_scenes], prec)
        tuned = autotune(rr, prec, lambda: bt.launch(graph=False), space={"steps_per_sync": (8, 16, 32, 48),
                                                                          "block_threads": (64, 256)}, reps=2)
        for _ in range(args.warmup):
            bt.launch(graph=True)
        rr.synchronize()
        t0 = time.perf_counter()
        kms = []
        for _ in range(args.steps):
            bt.launch(graph=True)
            kms.append(rr.last_kernel_ms())
        rr.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        util = rr.last_lane_utilisation()
        st_b, sm_b = rr.last_work()
        outs = bt.results()
        bt.close()
        rr.release(dbx)
        paths[name] = {"mrays_s": frames * n / dt / 1e6, "ms_per_launch": dt * 1e3, "kernel_ms": float(np.mean(kms)),
                       "frames": frames, "lane_utilisation": util, "tuning": tuned["best"], "rk2_steps": st_b,
                       "brick_samples": sm_b}
        if name == "batched_fp32":
            batch32 = outs
        log(f"[bench] {name}: {paths[name]}")
    # accuracy of the batched frames against the FP64 render of the same frames
    bt64 = Batch(rr, [(fs, db32) for fs in frame_scenes], PREC_FP64)
    bt64.launch(graph=False)
    outs64 = bt64.results()
    bt64.close()
    ok = np.concatenate([((a["status"] | b["status"]) & 8).ravel() == 0 for a, b in zip(outs64, batch32)])
    I64b = np.concatenate([o["iquv"][..., 0].ravel() for o in outs64])
    I32b = np.concatenate([o["iquv"][..., 0].ravel() for o in batch32]).astype(np.float64)
    paths["batched_fp32"]["rel_l1_I"] = rel_l1(I32b[ok], I64b[ok])

    log("[bench] 64-render batch: 128^2..1024^2, random spin and inclination, 4 bf16 bricks")
    bricks = [rr.upload(b, "bf16", "corner") for b in scenes.batch_bricks()]
    jobs = [(sc, bricks[k]) for sc, k in scenes.batch_scenes(64)]
    bt = Batch(rr, jobs, PREC_MIXED)
    tuned = autotune(rr, PREC_MIXED, lambda: bt.launch(graph=False), space={"steps_per_sync": (8, 16, 32, 48),
                                                                             "block_threads": (64, 256)}, reps=1)
    bt.launch(graph=True)
    rr.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        bt.launch(graph=True)
    rr.synchronize()
    dt = (time.perf_counter() - t0) / max(1, args.steps // 2)
    kms = rr.last_kernel_ms()
    util = rr.last_lane_utilisation()
    outs = bt.results()
    steps_all = np.concatenate([o["steps"].ravel() for o in outs])
    edges = [1, 64, 128, 256, 512, 1024, 2048, 4096, 4097]
    hist = np.histogram(steps_all, bins=edges)[0].tolist()
    rays_b = int(steps_all.size)
    paths["batch64_mixed_bf16"] = {"mrays_s": rays_b / dt / 1e6, "ms_per_launch": dt * 1e3, "kernel_ms": kms,
                                   "rays": rays_b, "lane_utilisation": util, "tuning": tuned["best"],
                                   "step_histogram": {"edges": edges, "counts": hist}}
    bt.close()
    for b in bricks:
        rr.release(b)
    # error of every render of the batch against the FP64 batch of the same scenes
    bricks64 = [rr.upload(b, "f32", "soa") for b in scenes.batch_bricks()]
    bt64 = Batch(rr, [(sc, bricks64[k]) for sc, k in scenes.batch_scenes(64)], PREC_FP64)
    bt64.launch(graph=False)
    outs64_b = bt64.results()
    bt64.close()
    for b in bricks64:
        rr.release(b)
    l1s, pix, quv, flag_diff, num, den = [], [], 0.0, 0, 0.0, 0.0
    for m, r in zip(outs, outs64_b):
        ok = ((m["status"] | r["status"]) & 8) == 0
        I, I0 = m["iquv"][..., 0].astype(np.float64), r["iquv"][..., 0]
        l1s.append(rel_l1(I[ok], I0[ok]))
        num += float(np.abs(I - I0)[ok].sum())
        den += float(np.abs(I0[ok]).sum())
        scale = float(np.abs(I0[ok]).max())
        big = ok & (np.abs(I0) > 1e-6 * scale)
        pix.append(np.abs(I - I0)[big] / np.abs(I0[big]))
        quv = max(quv, float(np.max(np.abs(m["iquv"][..., 1:].astype(np.float64) - r["iquv"][..., 1:])[ok])) / scale)
        flag_diff += int(np.sum(((m["status"] ^ r["status"]) & 8) != 0))
    pix = np.concatenate(pix)
    paths["batch64_mixed_bf16"].update({
        "rel_l1_I": num / den, "rel_l1_I_worst_render": max(l1s), "max_abs_QUV_over_max_I": quv,
        "pixel_rel_I": {"pixels": int(pix.size), "median": float(np.median(pix)), "p99": float(np.quantile(pix, 0.99)),
                        "max": float(pix.max())},
        "pixels_flagged_in_one_render_only": flag_diff})
    log(f"[bench] batch64: {paths['batch64_mixed_bf16']}")

    # ---- sampler at large N ----------------------------------------------------------------
    log(f"[bench] FP64 sampler, {args.sampler_points} points")
    paths["sampler_fp64"] = sampler_path(rr, brick, args.sampler_points, args.steps)
    log(f"[bench] sampler: {paths['sampler_fp64']}")

    log("[bench] sampler micro-benchmark: sample_and_step_stokes, 8x6x8 brick, 64 points, 1 iteration")
    paths["sampler_micro_64"] = micro_path(400)
    log(f"[bench] sampler micro: {paths['sampler_micro_64']}")

    # ---- coefficient-brick precompute (the input side of the path) ------------------------
    dims = tuple(int(v) for v in args.precompute_grid.split("x"))
    log(f"[bench] coefficient precompute, torus {args.precompute_grid}")
    paths["coefficient_precompute"] = precompute_path(rr, dims, args.steps, 0 if args.no_cpu else 64)
    log(f"[bench] precompute: {paths['coefficient_precompute']}")

    # ---- brick files: the reference's npz schema, written, read back and uploaded ------------
    import tempfile

    log("[bench] brick file round trip (npz, 128x64x128x11 float32)")
    with tempfile.TemporaryDirectory() as tmp:
        fpath = Path(tmp) / "brick.npz"
        t0 = time.perf_counter()
        brick.save_npz(fpath)
        t1 = time.perf_counter()
        back = scenes.Brick.load_npz(fpath)
        t2 = time.perf_counter()
        dbf = rr.upload(back, "f32", "corner")
        rr.synchronize()
        t3 = time.perf_counter()
        rr.release(dbf)
        mb = brick.coeffs.nbytes / 1e6
        paths["brick_files"] = {"coeff_MB": mb, "file_MB": fpath.stat().st_size / 1e6,
                                "save_s": t1 - t0, "load_s": t2 - t1, "upload_corner_s": t3 - t2,
                                "round_trip_equal": bool(np.array_equal(back.coeffs, brick.coeffs))}
    log(f"[bench] brick files: {paths['brick_files']}")

    # ---- learned surrogate + mixed precision (+ batched) -----------------------------------
    if not args.no_surrogate:
        paths["surrogate_mixed_bf16"], paths["surrogate_batched"] = surrogate_path(rr, scene, brick, ref, I64, args,
                                                                                  frame_scenes, outs64)
        sbp = paths["surrogate_batched"]
        sbp["ratio_vs_plain_fp32"] = sbp["mrays_s"] / value
        # what the network takes off the RK2 integrator: the frames' RK2 steps with and without it
        sbp["rk2_steps_replaced_frac"] = 1.0 - sbp["fallback_rk2_steps"] / max(paths["batched_mixed_bf16"]["rk2_steps"], 1)

    # headline of the layered method: surrogate + mixed (+ batched) vs the plain FP32 render
    target_err = 1e-3
    sur = paths.get("surrogate_batched")
    compliant = {k: v for k, v in paths.items() if v.get("rel_l1_I") is not None and v["rel_l1_I"] <= target_err
                 and "mrays_s" in v}
    best_name = max(compliant, key=lambda k: compliant[k]["mrays_s"]) if compliant else None
    headline = {
        "target": {"ratio": 10.0, "rel_l1_I": target_err},
        "surrogate_ratio": (sur["mrays_s"] / value) if sur else None,
        "surrogate_rel_l1_I": sur["rel_l1_I"] if sur else None,
        "meets_target": bool(sur and sur["mrays_s"] / value >= 10.0 and sur["rel_l1_I"] <= target_err),
        "best_path_within_error": best_name,
        "best_ratio_within_error": (compliant[best_name]["mrays_s"] / value) if best_name else None,
    }
    log(f"[bench] headline: {headline}")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (FP64 NumPy oracle)")
        cpu = cpu_baseline(scene, brick, args.cpu_seconds)

    result = {
        "metric": "Mrays/s, 1024² Kerr Stokes render: surrogate+mixed vs plain RK2 at ≤1e-3 I err",
        "value": value,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"plain FP32 RK2 Kerr Stokes render, a=0.94, i=17deg, {scene.width}x{scene.height} "
                               f"over +-20M, h=0.05(r-r+), <=4096 steps, brick 128x64x128x11 f32 (thin disk), "
                               f"corner-packed layout",
                   "path": "plain_fp32", "rays": n},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": TRAFFIC_PLAIN_FP32 if args.size == 1024 else None,
                     "work": {"rk2_steps": steps_total, "brick_samples": samples_total, "F_geo": F_GEO,
                              "F_sample": F_SAMPLE, "kernel_ms": plain["kernel_ms"]}},
        "cpu_baseline": cpu,
        "error": {"plain_fp32": {"rel_l1_I": err32, "max_abs_QUV": quv32,
                                 "pole_flagged_frac": float(1 - ok_plain.mean())}},
        "headline": headline,
        "paths": paths,
        "host": platform.node(),
    }
    print(json.dumps(result), flush=True)
    rr.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
