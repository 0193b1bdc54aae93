"""Learning curve of the MLP surrogate on the 1024^2 Kerr Stokes scene.

    python3 surrogate_curve.py [--seconds 90] [--knot-seconds 30] > curve.jsonl

Trains the surrogate (4 -> 4 x 256 SiLU -> 4 n_knots + 3, BF16 MFMA, Adam lr 1e-3, batch 8192) on 2^20
seeded RK2 geodesics (a ~ U[0, 0.998], i ~ U[5, 85] deg, alpha, beta ~ U[-20, 20] M) for ``--seconds`` of
training, and at log-spaced Adam steps prints one JSON line with:

* the held-out loss (2^14 other seeded rays);
* the per-knot position error on the held-out rays in cell units of the 128x64x128 brick: |dr| / (cell
  width at r), |d theta| / (pi / 63), |d phi| / (2 pi / 128), median and 90th percentile of the largest
  of the three over the recorded knots of unflagged rays;
* the surrogate render of the 1024^2 scene (a=0.94, i=17 deg, bf16 corner brick, mixed-precision RK2
  fallback): relative L1 error of I against the FP64 render over pixels flagged in neither, the fallback
  fraction and the render's kernel time.

Then trains 64- and 256-knot networks for ``--knot-seconds`` each and prints their final line.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "blackhole-sim_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

import numpy as np  # noqa: E402

from bhgrrt import scenes  # noqa: E402
from bhgrrt import surrogate as bsur  # noqa: E402
from bhgrrt.render import PREC_FP64, PREC_MIXED, Renderer  # noqa: E402


def knot_errors(sur: bsur.Surrogate, held: dict, brick: scenes.Brick) -> dict:
    """Largest of |dr|, |d theta|, |d phi| in cell units per recorded knot of the unflagged held-out rays."""
    c = sur.cfg
    n = int(held["n"])
    d_y = sur.r.alloc(n * c.n_out * 4)
    sur.mlp.forward(held["x"], n, d_y, transposed=False)
    sur.r.synchronize()
    y = d_y.to_host(np.empty((n, c.n_out), np.float32)).astype(np.float64)
    t = held["t"].to_host(np.empty((n, c.n_out), np.float32)).astype(np.float64)
    w = held["w"].to_host(np.empty(n, np.float32))
    d_y.free()
    nm = sur.norm
    mul, add = nm.out_mul.astype(np.float64), nm.out_add.astype(np.float64)
    y, t = y * mul + add, t * mul + add
    K = c.n_knots
    count = np.clip(np.rint(t[:, 4 * K + 2]).astype(int), 0, K)
    ok_ray = w > 0.5
    j = np.arange(K)[None, :]
    valid = ok_ray[:, None] & (j < count[:, None])
    tr, tt, tp = t[:, 0:4 * K:4], t[:, 1:4 * K:4], t[:, 2:4 * K:4]
    yr, yt, yp = y[:, 0:4 * K:4], y[:, 1:4 * K:4], y[:, 2:4 * K:4]
    r = brick.r.astype(np.float64)
    q = (r[-1] / r[0]) ** (1.0 / (r.size - 1))
    dr_cell = np.maximum(tr, r[0]) * (q - 1.0)
    e = np.maximum.reduce([np.abs(yr - tr) / dr_cell, np.abs(yt - tt) / (math.pi / (brick.theta.size - 1)),
                           np.abs(yp - tp) / (2.0 * math.pi / brick.phi.size)])[valid]
    return {"knot_err_cells_median": float(np.median(e)), "knot_err_cells_p90": float(np.quantile(e, 0.9)),
            "knots_compared": int(e.size)}


def image_error(sur: bsur.Surrogate, scene, dbx, ref) -> dict:
    out = sur.render(scene, dbx)
    ok = ((ref["status"] | out["status"]) & 8) == 0
    I, I0 = out["iquv"][..., 0].astype(np.float64)[ok], ref["iquv"][..., 0][ok]
    return {"image_rel_l1_I": float(np.abs(I - I0).sum() / np.abs(I0).sum()),
            "fallback_frac": out["fallback"] / scene.rays, "render_ms": out["ms"]}


def run(rr, brick, scene, dbx, ref, n_knots: int, seconds: float, checkpoints: bool) -> None:
    cfg = bsur.SurrogateConfig(n_knots=n_knots, fallback_precision=PREC_MIXED)
    sur = bsur.Surrogate(rr, brick, cfg, seed=0)
    data = sur.generate(1 << 20, 20240610)
    held = sur.generate(1 << 14, 20240611)
    batch = 8192
    nb = int(data["n"]) // batch
    marks = [2 ** k for k in range(6, 40)]
    done, t_train = 0, 0.0
    for mark in marks:
        if t_train >= seconds:
            break
        if not checkpoints and t_train + 1e-9 < seconds:
            mark = 1 << 40   # train to the time budget, measure once
        t0 = time.perf_counter()
        info = sur.train(data, mark - done, batch=batch, lr=1e-3, log_every=256, time_limit_s=seconds - t_train,
                         first_batch=done % nb)
        t_train += time.perf_counter() - t0
        done += info["steps"]
        line = {"n_knots": n_knots, "adam_steps": done, "train_s": round(t_train, 2),
                "epochs": round(done * batch / data["n"], 2), "flagged_train_frac": data["flagged"],
                "train_loss": info["loss"], "heldout_loss": sur.evaluate(held)}
        line.update(knot_errors(sur, held, brick))
        line.update(image_error(sur, scene, dbx, ref))
        print(json.dumps(line), flush=True)
    for k in ("x", "t", "w"):
        data[k].free()
        held[k].free()
    sur.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--seconds", type=float, default=90.0)
    ap.add_argument("--knot-seconds", type=float, default=30.0)
    ap.add_argument("--knots", default="64,256")
    args = ap.parse_args(argv)
    rr = Renderer(0)
    brick, scene = scenes.scale_brick(), scenes.scale_plain_scene()
    d64 = rr.upload(brick, "f32", "soa")
    ref = rr.render(scene, d64, PREC_FP64)
    rr.release(d64)
    dbx = rr.upload(brick, "bf16", "corner")
    print(json.dumps({"scene": "a=0.94, i=17 deg, 1024x1024 over +-20 M, h=0.05(r-r+), brick 128x64x128 bf16 corner",
                      "fp64_reference": "GPU FP64 render"}), flush=True)
    run(rr, brick, scene, dbx, ref, 128, args.seconds, True)
    for k in (int(v) for v in args.knots.split(",") if v):
        run(rr, brick, scene, dbx, ref, k, args.knot_seconds, False)
    rr.release(dbx)
    rr.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
