"""GPU parity of the exact Stokes step (PolarizedTransferConfig.use_exact_matrix_step,
polarized_transfer.py:51-80) against the FP64 oracle, which takes scipy's expm of the augmented
matrix like the reference.

* The small plain scene (32x18, fixed h = 20) and the scale scene at 48^2 in FP64: statuses and step counts
  equal; on the small scene relative I error <= 1e-8 on every pixel the oracle shows to be well conditioned at FP64 rounding
  (tests/conditioning.py), on the scale scene <= 1e-9 on 99 % of unflagged pixels.
* FP32 and mixed-precision exact renders against the FP64 exact render: rel-L1(I) <= 1e-4 / 1e-3.
* A batch of exact renders equals the single renders; the exact image differs from the RK2 one.
* The surrogate's knot transfer takes the RK2 step only and says so.
"""

from dataclasses import replace

import numpy as np
import pytest

from bhgrrt import scenes
from bhgrrt.render import PREC_FP32, PREC_FP64, PREC_MIXED, Batch
from conditioning import well_conditioned
from oracle import render as orender

pytestmark = pytest.mark.gpu


def rel_l1(a, b):
    return float(np.sum(np.abs(a - b)) / max(np.sum(np.abs(b)), 1e-300))


def test_small_scene_exact_fp64_matches_oracle(renderer):
    sc, br = replace(scenes.small_plain_scene(), exact_step=True), scenes.small_brick()
    db = renderer.upload(br, "f32")
    try:
        out = renderer.render(sc, db, PREC_FP64)
    finally:
        renderer.release(db)
    ref = orender.render_image(sc, br)
    np.testing.assert_array_equal(out["status"], ref["status"])
    np.testing.assert_array_equal(out["steps"], ref["steps"])
    I, I0 = out["iquv"][..., 0], ref["iquv"][..., 0]
    rel = np.abs(I - I0) / np.maximum(np.abs(I0), 1e-300)
    cond = well_conditioned(sc, br, 1e-12, 1e-6, ref)
    print(f"small scene exact fp64: {cond.mean():.3f} well conditioned, max rel there {rel[cond].max():.2e}; all pixels "
          f"median {np.median(rel):.2e} max {rel.max():.2e}")
    assert cond.mean() >= 0.5
    assert rel[cond].max() <= 1e-8


@pytest.fixture(scope="module")
def scale_exact(renderer):
    sc, br = replace(scenes.scale_plain_scene(48, 48), exact_step=True), scenes.scale_brick()
    db = renderer.upload(br, "f32")
    out64 = renderer.render(sc, db, PREC_FP64)
    yield sc, br, db, out64
    renderer.release(db)


def test_scale_scene_exact_fp64_matches_oracle(scale_exact):
    sc, br, db, out64 = scale_exact
    ref = orender.render_image(sc, br)
    ok = ((out64["status"] | ref["status"]) & orender.ST_POLE_CLAMPED) == 0
    same = (out64["status"] == ref["status"]) & (out64["steps"] == ref["steps"])
    assert same[ok].mean() >= 0.99
    I, I0 = out64["iquv"][..., 0][ok & same], ref["iquv"][..., 0][ok & same]
    rel = np.abs(I - I0) / np.maximum(np.abs(I0), 1e-300)
    print("48^2 exact fp64: median rel", np.median(rel), "p99", np.quantile(rel, 0.99))
    assert np.quantile(rel, 0.99) <= 1e-9


@pytest.mark.parametrize("storage,layout,precision,bar", [("f32", "corner", PREC_FP32, 1e-4),
                                                         ("bf16", "corner", PREC_MIXED, 1e-3)])
def test_scale_scene_exact_reduced_precision(renderer, scale_exact, storage, layout, precision, bar):
    sc, br, _, out64 = scale_exact
    db = renderer.upload(br, storage, layout)
    try:
        out = renderer.render(sc, db, precision)
        rk2 = renderer.render(replace(sc, exact_step=False), db, precision)
    finally:
        renderer.release(db)
    ok = ((out64["status"] | out["status"]) & orender.ST_POLE_CLAMPED) == 0
    e = rel_l1(out["iquv"][..., 0][ok].astype(np.float64), out64["iquv"][..., 0][ok])
    d = rel_l1(rk2["iquv"][..., 0][ok].astype(np.float64), out["iquv"][..., 0][ok].astype(np.float64))
    print(f"exact {storage}/{precision}: rel L1 vs FP64 exact {e:.2e}; RK2 vs exact {d:.2e}")
    assert e <= bar
    assert d > 0.0


def test_exact_batch_equals_single_renders(renderer):
    bricks = scenes.batch_bricks()[:2]
    dbs = [renderer.upload(b, "f32", "corner") for b in bricks]
    jobs = [(replace(sc.with_size(w, w), exact_step=True), dbs[k % 2]) for k, (sc, _), w in
            zip(range(3), scenes.batch_scenes(3, seed=5), (24, 40, 17))]
    batch = Batch(renderer, jobs, PREC_FP32)
    saved = renderer.get_tuning(PREC_FP32)
    # the batched launch has no two-wave split of lone rays: compare with single renders without it
    renderer.set_tuning(PREC_FP32, coop_steps=0)
    try:
        batch.launch(graph=False)
        for (sc, db), g in zip(jobs, batch.results()):
            ref = renderer.render(sc, db, PREC_FP32)
            np.testing.assert_array_equal(g["status"], ref["status"])
            np.testing.assert_array_equal(g["steps"], ref["steps"])
            np.testing.assert_array_equal(g["iquv"], ref["iquv"])
    finally:
        renderer.set_tuning(PREC_FP32, **saved)
        batch.close()
        for d in dbs:
            renderer.release(d)


def test_surrogate_transfer_rejects_the_exact_step(renderer):
    from bhgrrt import surrogate as bsur

    br = scenes.scale_brick()
    cfg = bsur.SurrogateConfig(n_knots=8, width=256, depth=2)
    sur = bsur.Surrogate(renderer, br, cfg, seed=0)
    n_out = cfg.n_out
    sur.norm = bsur.Normalisation(in_mul=np.ones(4), in_add=np.zeros(4), out_mul=np.ones(n_out), out_add=np.zeros(n_out))
    db = renderer.upload(br, "bf16", "corner")
    sc = replace(scenes.scale_plain_scene(16, 16), exact_step=True)
    n = sc.rays
    bufs = (renderer.alloc(n * 16), renderer.alloc(n * 4), renderer.alloc(n))
    try:
        with pytest.raises(ValueError, match="RK2 Stokes step only"):
            sur.render_device(sc, db, *bufs)
    finally:
        for b in bufs:
            b.free()
        renderer.release(db)
        sur.close()
