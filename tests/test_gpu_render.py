"""GPU parity of the fused render (FP64 and plain FP32) against the FP64 oracle.

* Small plain scene (a=0.9, r_obs=1000, i=60 deg, 32x18, 64 fixed steps of h=20,
  8x6x8 brick): every pixel against the oracle.
* Scale scene (a=0.94, i=17 deg, 1024^2, h = 0.05 (r - r_+), <= 4096 steps,
  128x64x128 disk brick): a seeded pixel subset against the oracle, plus
  size-independent checks on the full image (FP32 vs FP64 agreement, step and
  status statistics).

Rays flagged by the theta clamp (status bit 3) left the geodesic at a pole, so
their Stokes values are excluded from error statistics (counted instead).  In
the small scene, 64 fixed steps of h = 20 from r = 1000 overstep a pole on
almost every ray (and the explicit Stokes step is unstable where alpha ds > 2),
so there every pixel is compared: statuses and step counts exactly, values to
the accuracy the amplification allows.
Tolerances: FP64 GPU vs FP64 oracle -- identical status and step count,
per-pixel relative I error <= 1e-9 on 99% of unflagged pixels of the scale
scene; plain FP32 vs FP64 -- relative L1 error on I <= 2e-3 over unflagged
pixels of the scale scene.
"""

import numpy as np
import pytest

from bhgrrt import scenes
from bhgrrt.render import PREC_FP32, PREC_FP64, PREC_MIXED
from oracle import render as orender

pytestmark = pytest.mark.gpu


def rel_l1(a, b):
    return float(np.sum(np.abs(a - b)) / max(np.sum(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def small(renderer):
    sc, br = scenes.small_plain_scene(), scenes.small_brick()
    db = renderer.upload(br, "f32")
    ref = orender.render_image(sc, br)
    yield sc, br, db, ref
    renderer.release(db)


def test_small_scene_fp64_matches_oracle(renderer, small):
    sc, br, db, ref = small
    out = renderer.render(sc, db, PREC_FP64)
    np.testing.assert_array_equal(out["status"], ref["status"])
    np.testing.assert_array_equal(out["steps"], ref["steps"])
    I, I0 = out["iquv"][..., 0], ref["iquv"][..., 0]
    rel = np.abs(I - I0) / np.maximum(np.abs(I0), 1e-300)
    print("small fp64: median rel", np.median(rel), "p90", np.quantile(rel, 0.9), "max", rel.max())
    assert np.median(rel) <= 1e-10


def test_small_scene_fp32_runs_the_same_schedule(renderer, small):
    """h = 20 makes every FP32 ray chaotic here, so only schedule-level properties are compared."""
    sc, br, db, ref = small
    out = renderer.render(sc, db, PREC_FP32)
    assert np.all(np.isfinite(out["iquv"]))
    assert out["steps"].min() >= 1 and out["steps"].max() <= sc.max_steps
    agree = (out["status"] & 7) == (ref["status"] & 7)
    print("small fp32: termination agreement", agree.mean(), "steps agreement", (out["steps"] == ref["steps"]).mean())
    assert agree.mean() >= 0.3


def test_scale_scene_fp32_matches_oracle_at_48(renderer):
    """The 1024^2 scene's camera and brick at 48x48: plain FP32 (corner layout) against the FP64 oracle."""
    sc, br = scenes.scale_plain_scene(48, 48), scenes.scale_brick()
    ref = orender.render_image(sc, br)
    db = renderer.upload(br, "f32", "corner")
    try:
        out = renderer.render(sc, db, PREC_FP32)
    finally:
        renderer.release(db)
    ok = ((ref["status"] | out["status"]) & orender.ST_POLE_CLAMPED) == 0
    assert np.mean((out["status"] == ref["status"])[ok]) >= 0.99
    e = rel_l1(out["iquv"][..., 0][ok].astype(np.float64), ref["iquv"][..., 0][ok])
    print("48^2 fp32 vs oracle rel L1 I", e)
    assert e <= 2e-3


@pytest.fixture(scope="module")
def scale(renderer):
    sc, br = scenes.scale_plain_scene(), scenes.scale_brick()
    db = renderer.upload(br, "f32")
    out64 = renderer.render(sc, db, PREC_FP64)
    out32 = renderer.render(sc, db, PREC_FP32)
    yield sc, br, db, out64, out32
    renderer.release(db)


def test_scale_scene_fp64_matches_oracle_on_pixel_subset(scale):
    sc, br, db, out64, _ = scale
    rng = np.random.default_rng(12345)
    pix = np.sort(rng.choice(sc.width * sc.height, 3000, replace=False))
    ref = orender.render_rays(sc, br, pix)
    st = out64["status"].ravel()[pix]
    steps = out64["steps"].ravel()[pix]
    I = out64["iquv"].reshape(-1, 4)[pix]
    ok = (ref["status"] & orender.ST_POLE_CLAMPED) == 0
    same = (st == ref["status"]) & (steps == ref["steps"])
    print("scale fp64: status+steps agreement", same.mean(), "flagged", 1 - ok.mean())
    assert same[ok].mean() >= 0.99
    rel = np.abs(I[:, 0] - ref["iquv"][:, 0]) / np.maximum(np.abs(ref["iquv"][:, 0]), 1e-300)
    print("scale fp64: median rel", np.median(rel[ok]), "p99", np.quantile(rel[ok], 0.99))
    assert np.quantile(rel[ok & same], 0.99) <= 1e-9
    assert rel_l1(I[ok, 0], ref["iquv"][ok, 0]) <= 1e-7


def test_scale_scene_fp32_error_against_fp64(scale):
    sc, br, db, out64, out32 = scale
    ok = ((out64["status"] | out32["status"]) & orender.ST_POLE_CLAMPED) == 0
    I64 = out64["iquv"][..., 0][ok]
    I32 = out32["iquv"][..., 0][ok].astype(np.float64)
    e = rel_l1(I32, I64)
    qu = np.max(np.abs(out32["iquv"][..., 1:][ok] - out64["iquv"][..., 1:][ok]))
    print("scale fp32 vs fp64: rel L1 I", e, "max abs QUV", qu, "flagged frac", 1 - ok.mean())
    assert e <= 2e-3
    steps = out32["steps"]
    assert steps.min() >= 1 and steps.max() <= sc.max_steps
    assert 100 < steps.mean() < 400


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_scale_scene_mixed_precision_error_against_fp64(renderer, scale, storage):
    """BF16/FP16 brick, FP32 geodesic with FP64 islands: relative L1 on I <= 1e-3 vs the FP64 render."""
    sc, br, db, out64, _ = scale
    dbx = renderer.upload(br, storage, "corner")
    try:
        out = renderer.render(sc, dbx, PREC_MIXED)
    finally:
        renderer.release(dbx)
    ok = ((out64["status"] | out["status"]) & orender.ST_POLE_CLAMPED) == 0
    e = rel_l1(out["iquv"][..., 0][ok].astype(np.float64), out64["iquv"][..., 0][ok])
    qu = np.max(np.abs(out["iquv"][..., 1:][ok] - out64["iquv"][..., 1:][ok]))
    print(f"scale mixed {storage} vs fp64: rel L1 I {e:.3e} max abs QUV {qu:.3e} flagged {1 - ok.mean():.4f}")
    assert e <= 1e-3


@pytest.mark.parametrize("storage,precision", [("f32", PREC_FP32), ("bf16", PREC_FP32), ("bf16", PREC_MIXED),
                                               ("f16", PREC_MIXED)])
def test_corner_layout_matches_soa(renderer, storage, precision):
    """Corner-packed cell records hold the same values as the SoA planes and are summed in the same order;
    only the compiler's multiply-add fusion may differ (1 ulp per operation)."""
    sc = scenes.scale_plain_scene(256, 256)
    br = scenes.scale_brick()
    a = renderer.upload(br, storage, "soa")
    b = renderer.upload(br, storage, "corner")
    try:
        oa = renderer.render(sc, a, precision)
        ob = renderer.render(sc, b, precision)
    finally:
        renderer.release(a)
        renderer.release(b)
    ok = ((oa["status"] | ob["status"]) & orender.ST_POLE_CLAMPED) == 0
    assert np.mean(oa["status"] == ob["status"]) >= 0.995
    assert rel_l1(ob["iquv"][..., 0][ok].astype(np.float64), oa["iquv"][..., 0][ok].astype(np.float64)) <= 1e-5


@pytest.mark.parametrize("storage,layout,precision", [("f32", "corner", PREC_FP32), ("bf16", "corner", PREC_MIXED),
                                                      ("f32", "soa", PREC_FP64)])
def test_batch_matches_individual_renders(renderer, storage, layout, precision):
    """One persistent launch over renders of different size, spin, inclination and brick gives each
    render's own result (statuses and step counts identical; values to FMA-fusion rounding)."""
    from bhgrrt.render import Batch

    bricks = scenes.batch_bricks()[:3]
    dbs = [renderer.upload(b, storage, layout) for b in bricks]
    jobs = [(sc.with_size(w, w), dbs[k % 3]) for k, (sc, _), w in
            zip(range(5), scenes.batch_scenes(5, seed=7), (40, 64, 24, 56, 33))]
    batch = Batch(renderer, jobs, precision)
    try:
        for graph in (False, True, True):      # direct launch, graph capture, graph replay
            batch.launch(graph=graph)
            got = batch.results()
            for (sc, db), g in zip(jobs, got):
                ref = renderer.render(sc, db, precision)
                assert np.mean(g["status"] == ref["status"]) >= 0.99
                assert np.mean(g["steps"] == ref["steps"]) >= 0.99
                ok = ((g["status"] | ref["status"]) & orender.ST_POLE_CLAMPED) == 0
                tol = 1e-12 if precision == PREC_FP64 else 1e-5
                assert rel_l1(g["iquv"][..., 0][ok].astype(np.float64), ref["iquv"][..., 0][ok].astype(np.float64)) <= tol
    finally:
        batch.close()
        for d in dbs:
            renderer.release(d)
