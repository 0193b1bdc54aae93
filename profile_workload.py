"""Fixed workload for rocprofv3 runs: the 1024^2 Kerr Stokes scene on every path, plus a calibration copy.

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- \\
        python3 profile_workload.py
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o run -- \\
        python3 profile_workload.py

Render paths are storage:layout:precision[:exact] (``exact`` selects the exact Stokes step); each is
rendered twice (the second launch is the measured one).  ``--surrogate`` adds the learned surrogate: FP64 geodesic knots for
``--surrogate-rays`` seeded rays, ``--train-steps`` Adam steps on the BF16 MFMA kernels and two
1024^2 surrogate renders (bf16 corner brick, mixed-precision fallback); ``--surrogate-frames N`` adds two
launches of the batched surrogate over N frames (camera phi = 2 pi k / N).  ``--sampler N`` adds two
launches of the FP64 trilinear sampler on N uniform points; ``--mlp-rows N`` times the surrogate
network's inference on N rows (fused kernel, then per-layer GEMMs); ``--precompute NRxNTxNP`` runs the
coefficient-brick precompute on the analytic torus.  A 256 MiB device-to-device copy comes
last, so the FETCH_SIZE / WRITE_SIZE counters can be scaled by what they read for a copy of known
size.
"""

from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import math
import sys
from pathlib import Path

_ROOT = Path(__file__).resolve().parent
for _p in (_ROOT, _ROOT / "blackhole-sim_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

import numpy as np  # noqa: E402

from bhgrrt import _lib, scenes  # noqa: E402
from bhgrrt.render import Renderer  # noqa: E402

PREC = {"fp64": 0, "fp32": 1, "mixed": 2}
COPY_BYTES = 256 << 20


def run_sampler(r: Renderer, brick, n: int) -> None:
    rng = np.random.default_rng(20240607)
    pts = np.stack([rng.uniform(float(brick.r[0]), float(brick.r[-1]), n),
                    rng.uniform(float(brick.theta[0]), float(brick.theta[-1]), n),
                    rng.uniform(0.0, 2.0 * math.pi, n)], axis=1)
    arrs = [np.ascontiguousarray(brick.coeffs, np.float64)] + [np.ascontiguousarray(g, np.float64)
                                                                for g in (brick.r, brick.theta, brick.phi)] + [pts]
    bufs = [r.alloc(a.nbytes) for a in arrs]
    for b, a in zip(bufs, arrs):
        b.from_host(a)
    out = r.alloc(n * 88)
    for _ in range(2):
        _lib.check(r._lib.bh_sample_brick_trilinear_device(r.ctx, bufs[0].ptr, bufs[1].ptr, brick.r.size, bufs[2].ptr,
                                                           brick.theta.size, bufs[3].ptr, brick.phi.size, bufs[4].ptr,
                                                           C.c_int64(n), out.ptr))
        r.synchronize()
    print(f"sampler: {n} points, {r.last_kernel_ms():.3f} ms", flush=True)
    for b in bufs + [out]:
        b.free()


def run_precompute(r: Renderer, dims: tuple[int, int, int]) -> None:
    from bhgrrt import coefficient_bricks as cb
    from bhgrrt.grmhd import generate_analytic_grmhd_torus

    snap = generate_analytic_grmhd_torus(spin_a=0.85, nr=dims[0], ntheta=dims[1], nphi=dims[2])
    sc = cb.PhysicalScaling.from_mdot(cb.BlackHoleSystem.m87_star(0.85), 1e-4)
    fields = [np.ascontiguousarray(a) for a in (snap.rho, snap.theta_e, snap.b_con)]
    bufs = [r.alloc(a.nbytes) for a in fields]
    for b, a in zip(bufs, fields):
        b.from_host(a)
    n = snap.rho.size
    out = r.alloc(n * 11 * 4)
    desc = cb.synchrotron_desc(sc, 230e9)
    for _ in range(2):
        _lib.check(r._lib.bh_precompute_coefficient_bricks_device(r.ctx, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                                                  C.c_int64(n), C.byref(desc), 0, out.ptr))
        r.synchronize()
    print(f"coefficient precompute: {n} cells, {r.last_kernel_ms():.3f} ms", flush=True)
    for b in bufs + [out]:
        b.free()


def run_mlp(r: Renderer, rows: int) -> None:
    from bhgrrt import surrogate as bsur

    mlp = bsur.MLP(r, 4, 256, 4, 515)
    mlp.set_params(mlp.init_params(3))
    x, y = r.alloc(rows * 16), r.alloc(rows * 515 * 4)
    x.from_host(np.random.default_rng(1).uniform(-1, 1, (rows, 4)).astype(np.float32))
    for fused in (True, False):
        mlp.set_fused(fused)
        ms = []
        for _ in range(4):
            mlp.forward(x, rows, y, transposed=True)
            r.synchronize()
            ms.append(r.last_kernel_ms())
        t = float(np.median(ms[1:]))
        print(f"mlp inference {'fused' if fused else 'per-layer'}: {rows} rows, {t:.3f} ms, "
              f"{2 * 3.31e5 * rows / t / 1e9:.1f} TFLOP/s", flush=True)
    x.free()
    y.free()
    mlp.close()


def run_surrogate(r: Renderer, sc, brick, rays: int, steps: int, frames: int) -> None:
    from bhgrrt import surrogate as bsur

    cfg = bsur.SurrogateConfig(fallback_precision=PREC["mixed"])
    sur = bsur.Surrogate(r, brick, cfg, seed=1)
    data = sur.generate(rays, 20240605)
    print(f"knots: {rays} rays, {data['ms']:.3f} ms", flush=True)
    info = sur.train(data, steps, log_every=max(steps, 1))
    print(f"training: {info['steps']} Adam steps, loss {info['loss']:.3f}", flush=True)
    for k in ("x", "t", "w"):
        data[k].free()
    db = r.upload(brick, "bf16", "corner")
    n = sc.rays
    bufs = (r.alloc(n * 16), r.alloc(n * 4), r.alloc(n))
    for _ in range(2):
        sur.render_device(sc, db, *bufs)
        r.synchronize()
    print(f"surrogate render: {r.last_kernel_ms():.3f} ms, fallback {sur.last_fallback()} rays", flush=True)
    for b in bufs:
        b.free()
    if frames:
        fs = [dataclasses.replace(sc, camera=dataclasses.replace(sc.camera, phi=2.0 * math.pi * k / frames))
              for k in range(frames)]
        sb = sur.batch([(f, db) for f in fs])
        for _ in range(2):
            sb.launch(graph=True)
            r.synchronize()
        print(f"surrogate batch of {frames} frames: {r.last_kernel_ms():.3f} ms, fallback {sb.fallback()[1]} rays",
              flush=True)
        sb.close()
    r.release(db)
    sur.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--paths", default="f32:corner:fp32,bf16:corner:mixed,f16:corner:mixed,f32:soa:fp32")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--surrogate", action="store_true")
    ap.add_argument("--surrogate-rays", type=int, default=1 << 18)
    ap.add_argument("--train-steps", type=int, default=64)
    ap.add_argument("--surrogate-frames", type=int, default=0, help="with --surrogate: also launch the batched "
                    "surrogate sequence (one captured graph) twice over this many frames of the scene")
    ap.add_argument("--sampler", type=int, default=0, help="points of the FP64 sampler launch (0: none)")
    ap.add_argument("--mlp-rows", type=int, default=0, help="time transposed inference of the 4->256x4->515 "
                    "network on this many rows, fused kernel and per-layer GEMMs (0: skip)")
    ap.add_argument("--latency-sizes", default="", help="e.g. 16,64,256,1024: plain FP32 renders of the scene at "
                    "these sizes, printing kernel time, longest ray and time per step of the longest ray")
    ap.add_argument("--precompute", default="", help="e.g. 128x64x128: two launches of the coefficient-brick "
                    "precompute on the analytic torus of that grid (float32 out)")
    args = ap.parse_args(argv)
    sc, br = scenes.scale_plain_scene(args.size, args.size), scenes.scale_brick()
    r = Renderer(0)
    n = sc.rays
    bufs = (r.alloc(n * 32), r.alloc(n * 4), r.alloc(n))
    for spec in [p for p in args.paths.split(",") if p]:
        storage, layout, prec, *rule = spec.split(":")
        scx = dataclasses.replace(sc, exact_step=rule == ["exact"])
        db = r.upload(br, storage, layout)
        for _ in range(2):
            r.render_device(scx, db, PREC[prec], *bufs)
            r.synchronize()
        steps, samples = r.last_work()
        print(f"{spec}: {r.last_kernel_ms():.3f} ms, {steps} RK2 steps, {samples} brick samples", flush=True)
        r.release(db)
    for b in bufs:
        b.free()
    if args.latency_sizes:
        db = r.upload(br, "f32", "corner")
        for size in [int(v) for v in args.latency_sizes.split(",")]:
            scn = scenes.scale_plain_scene(size, size)
            ms = []
            for _ in range(3):
                out = r.render(scn, db, PREC["fp32"])
                ms.append(r.last_kernel_ms())
            t = float(np.median(ms))
            longest = int(out["steps"].max())
            print(f"latency {size}x{size}: kernel {t:.3f} ms, longest ray {longest} steps, "
                  f"{1e3 * t / longest:.2f} us per step of the longest ray, {int(out['steps'].sum())} steps in all",
                  flush=True)
        r.release(db)
    if args.precompute:
        run_precompute(r, tuple(int(v) for v in args.precompute.split("x")))
    if args.mlp_rows:
        run_mlp(r, args.mlp_rows)
    if args.sampler:
        run_sampler(r, br, args.sampler)
    if args.surrogate:
        run_surrogate(r, sc, br, args.surrogate_rays, args.train_steps, args.surrogate_frames)
    a, b = r.alloc(COPY_BYTES), r.alloc(COPY_BYTES)
    _lib.check(_lib.lib().bh_copy_device(r.ctx, b.ptr, a.ptr, COPY_BYTES))
    print(f"calibration copy: {COPY_BYTES} bytes", flush=True)
    r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
