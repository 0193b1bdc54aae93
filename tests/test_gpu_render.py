"""GPU parity of the fused render (FP64 and plain FP32) against the FP64 oracle.

* Small plain scene (a=0.9, r_obs=1000, i=60 deg, 32x18, 64 fixed RK2 steps of h=20, 8x6x8 brick).  Steps of 20 M jump across the coordinate axis near the hole, so every ray is pole-flagged and
  most of them amplify rounding errors enormously (a 1-ulp change of the camera moves I by 38 % at the
  90th percentile in the oracle itself).  Bars: statuses and step counts equal to the oracle on every
  pixel (FP64); I within 1e-9 relative on every pixel the oracle shows to be well conditioned at FP64
  rounding (I moves < 1e-6 under a 1e-12 change of the camera; 72 % of the pixels; measured 7e-11);
  FP32 against the FP64 render on the pixels well conditioned at FP32 rounding (I moves < 1e-2 under a
  1e-6 change; 64 %): statuses and steps on >= 45 % of them (measured 56 %: FP32 rounding inside a
  step that crosses the axis is not a perturbation of the camera) and I within 5e-4 on every one of
  those that agree (measured 4.7e-5).
* Scale scene (a=0.94, i=17 deg, 1024^2, h = 0.05 (r - r_+), <= 4096 steps, 128x64x128 disk brick): a seeded pixel subset of the FP64 render against the oracle (statuses and steps on >= 99 % of
  unflagged pixels, I within 1e-9 on 99 % of those that agree); the full FP32 image against the FP64
  render: relative L1 error of I <= 1e-4, per-pixel relative I error and max(|dQ|, |dU|, |dV|) / |I|
  <= 1e-4 at the 99th percentile, max |dQUV| <= 3e-2 of the image's largest |I| (a few near-critical
  rays diverge: the oracle confirms the FP64 values of the worst pixels), at most 0.1 % of the pixels
  flagged in one render only.  Mixed precision (bf16 / f16 bricks, FP64 islands) against the FP64 render: relative L1 error
  of I <= 1e-3 (bf16) / 2e-4 (f16), per-pixel p99 of I and Q/U/V <= 3e-3 / 5e-4, the same max and flag
  bars.
* Rays flagged by the theta clamp (status bit 3) left the geodesic at a pole: on the scale scene their
  Stokes values are excluded from error statistics and counted.
"""

import numpy as np
import pytest

from bhgrrt import scenes
from bhgrrt.render import PREC_FP32, PREC_FP64, PREC_MIXED
from conditioning import well_conditioned
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
    cond = well_conditioned(sc, br, 1e-12, 1e-6, ref)
    print(f"small fp64: {cond.mean():.3f} well conditioned, max rel there {rel[cond].max():.2e}; all pixels median "
          f"{np.median(rel):.2e} p90 {np.quantile(rel, 0.9):.2e} max {rel.max():.2e}")
    assert cond.mean() >= 0.5
    assert rel[cond].max() <= 1e-9


def test_small_scene_fp32_against_fp64_on_conditioned_rays(renderer, small):
    sc, br, db, ref = small
    o32 = renderer.render(sc, db, PREC_FP32)
    o64 = renderer.render(sc, db, PREC_FP64)
    assert np.all(np.isfinite(o32["iquv"]))
    assert o32["steps"].min() >= 1 and o32["steps"].max() <= sc.max_steps
    cond = well_conditioned(sc, br, 1e-6, 1e-2, ref)
    same = (o32["status"] == o64["status"]) & (o32["steps"] == o64["steps"])
    I, I0 = o32["iquv"][..., 0].astype(np.float64), o64["iquv"][..., 0]
    rel = np.abs(I - I0) / np.maximum(np.abs(I0), 1e-300)
    agree = cond & same
    print(f"small fp32: {cond.mean():.3f} well conditioned at FP32 rounding, statuses+steps there "
          f"{same[cond].mean():.3f}, rel I p95 {np.quantile(rel[agree], 0.95):.2e} max {rel[agree].max():.2e}; "
          f"all pixels: termination agreement {((o32['status'] & 7) == (o64['status'] & 7)).mean():.3f}")
    assert cond.mean() >= 0.4
    assert same[cond].mean() >= 0.45
    assert rel[agree].max() <= 5e-4


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
    assert e <= 1e-4


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


def _image_errors(out, out64):
    """Errors of a render against the FP64 render over pixels flagged in neither: relative L1 of I,
    per-pixel relative I (pixels with |I| above 1e-6 of the largest), max |dQUV| / max |I|, and the number
    of pixels flagged in one render only."""
    ok = ((out64["status"] | out["status"]) & orender.ST_POLE_CLAMPED) == 0
    I, I64 = out["iquv"][..., 0].astype(np.float64), out64["iquv"][..., 0]
    scale = np.abs(I64[ok]).max()
    big = ok & (np.abs(I64) > 1e-6 * scale)
    pix = np.abs(I - I64)[big] / np.abs(I64[big])
    dquv = np.max(np.abs(out["iquv"][..., 1:].astype(np.float64) - out64["iquv"][..., 1:]), axis=-1)
    quv = float(dquv[ok].max()) / scale
    pix_quv = dquv[big] / np.abs(I64[big])
    flag_diff = int(np.sum(((out["status"] ^ out64["status"]) & orender.ST_POLE_CLAMPED) != 0))
    return rel_l1(I[ok], I64[ok]), pix, quv, flag_diff, ok, pix_quv


def test_scale_scene_fp32_error_against_fp64(scale):
    sc, br, db, out64, out32 = scale
    e, pix, quv, flag_diff, ok, pix_quv = _image_errors(out32, out64)
    print(f"scale fp32 vs fp64: rel L1 I {e:.2e}, pixel rel I p99 {np.quantile(pix, 0.99):.2e} max {pix.max():.2e}, "
          f"pixel |dQUV| / |I| p99 {np.quantile(pix_quv, 0.99):.2e}, max |dQUV| / max|I| {quv:.2e}, "
          f"flagged {1 - ok.mean():.4f}, flagged in one render only {flag_diff}")
    assert e <= 1e-4
    assert np.quantile(pix, 0.99) <= 1e-4
    assert np.quantile(pix_quv, 0.99) <= 1e-4
    assert quv <= 3e-2
    assert flag_diff <= 0.001 * ok.size
    # the pixels where FP32 is furthest from FP64: the FP64 render agrees with the oracle there, so the
    # outliers are FP32 rays that diverged (near-critical orbits amplify their rounding)
    I32, I64 = out32["iquv"][..., 0].astype(np.float64).ravel(), out64["iquv"][..., 0].ravel()
    okf = ok.ravel() & (np.abs(I64) > 1e-6 * np.abs(I64[ok.ravel()]).max())
    cand = np.flatnonzero(okf)
    worst = cand[np.argsort(np.abs(I32 - I64)[cand] / np.abs(I64[cand]))[-4:]]
    o = orender.render_rays(sc, br, worst)
    print("worst fp32 pixels:", worst.tolist(), "steps", out64["steps"].ravel()[worst].tolist(), "fp32 rel",
          (np.abs(I32 - I64)[worst] / np.abs(I64[worst])).tolist())
    np.testing.assert_array_equal(o["steps"], out64["steps"].ravel()[worst])
    np.testing.assert_allclose(o["iquv"][:, 0], I64[worst], rtol=1e-9, atol=0)
    steps = out32["steps"]
    assert steps.min() >= 1 and steps.max() <= sc.max_steps
    assert 100 < steps.mean() < 400


@pytest.mark.parametrize("storage,l1_bar,p99_bar", [("bf16", 1e-3, 3e-3), ("f16", 2e-4, 5e-4)])
def test_scale_scene_mixed_precision_error_against_fp64(renderer, scale, storage, l1_bar, p99_bar):
    """BF16/FP16 brick, FP32 geodesic with FP64 islands, against the FP64 render."""
    sc, br, db, out64, _ = scale
    dbx = renderer.upload(br, storage, "corner")
    try:
        out = renderer.render(sc, dbx, PREC_MIXED)
    finally:
        renderer.release(dbx)
    e, pix, quv, flag_diff, ok, pix_quv = _image_errors(out, out64)
    print(f"scale mixed {storage} vs fp64: rel L1 I {e:.3e}, pixel rel I p99 {np.quantile(pix, 0.99):.2e} max "
          f"{pix.max():.2e}, pixel |dQUV| / |I| p99 {np.quantile(pix_quv, 0.99):.2e}, max |dQUV| / max|I| "
          f"{quv:.2e}, flagged {1 - ok.mean():.4f}, one render only {flag_diff}")
    assert e <= l1_bar
    assert np.quantile(pix, 0.99) <= p99_bar
    assert np.quantile(pix_quv, 0.99) <= p99_bar
    assert quv <= 3e-2
    assert flag_diff <= 0.001 * ok.size


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


@pytest.mark.parametrize("storage,precision", [("f32", PREC_FP32), ("bf16", PREC_MIXED)])
def test_corner_brick_beyond_4_gib(storage, precision):
    """Corner records of 2 GiB and more are read through 64-bit addresses: a 256 x 192 x 256 brick (4.39 GB
    of FP32 corner records, past what a 32-bit buffer offset reaches; 2.19 GB of bf16 records, channel-major)
    renders like its SoA copy.  Rays from the camera enter the brick through its outermost radial cells,
    whose records lie at the end of the buffer."""
    from bhgrrt.render import Renderer

    print(f"corner brick beyond 2 GiB ({storage}): building 256x192x256", flush=True)
    br = scenes.make_brick(256, 192, 256, spin=0.94, seed=7, disk=True)
    cells = 255 * 191 * 256
    assert cells * 8 * 11 * (4 if storage == "f32" else 2) > (1 << 31)
    sc = scenes.scale_plain_scene(128, 128)
    rr = Renderer(0)
    try:
        a = rr.upload(br, storage, "soa")
        b = rr.upload(br, storage, "corner")
        print(f"corner brick beyond 2 GiB ({storage}): uploaded, rendering", flush=True)
        oa = rr.render(sc, a, precision)
        ob = rr.render(sc, b, precision)
        _, samples = rr.last_work()
    finally:
        rr.close()
    ok = ((oa["status"] | ob["status"]) & orender.ST_POLE_CLAMPED) == 0
    l1 = rel_l1(ob["iquv"][..., 0][ok].astype(np.float64), oa["iquv"][..., 0][ok].astype(np.float64))
    print(f"corner brick beyond 2 GiB ({storage}): {samples} samples, statuses equal on "
          f"{np.mean(oa['status'] == ob['status']):.4f}, rel L1 I {l1:.2e}")
    assert samples > 0
    assert np.mean(oa["status"] == ob["status"]) >= 0.995
    assert l1 <= 1e-5


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
