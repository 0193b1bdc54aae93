"""GPU render edge cases, schedule invariance and argument errors.

* Launch shape and schedule (block size, persistent grid, steps between refills, refill
  threshold, queue count, tile order) decide only which lane traces which ray, so every
  tuning gives bit-identical images, step counts and statuses -- also for the batched
  launch, before and after ``autotune`` picks its tuning.
* Ragged and degenerate image sizes (1x1, one row, one column, sizes that leave 8x8 tiles
  partly outside the image) and extreme scenes (Schwarzschild, a = 0.998, observers 2 deg
  from the pole and edge-on, camera azimuth != 0, retrograde spin) rendered in FP64 against
  the FP64 oracle: statuses and step counts identical on >= 98 % of unflagged rays (all of
  them below 100 rays), relative I error <= 1e-12 on 95 % of those and <= 1e-9 on all,
  Q/U/V within 1e-9 of the brightest I.  The GPU evaluates the Kerr equations in separated (Carter) form and the
  oracle in matrix form; the FP64 rounding differences this leaves grow only on rays that
  skim the photon orbit.
* Invalid arguments are ``ValueError`` / ``RuntimeError`` with the library's message.
"""

import math
from dataclasses import replace

import numpy as np
import pytest

from bhgrrt import scenes
from bhgrrt.render import PREC_FP32, PREC_FP64, PREC_MIXED, Batch
from oracle import render as orender

pytestmark = pytest.mark.gpu

# coop_steps=0: the lone-ray path without the two-wave split (the split is checked against it below)
TUNINGS = [
    dict(block_threads=64, blocks_per_cu=1, steps_per_sync=1, refill_min=1, queues=1, order=0, coop_steps=0),
    dict(block_threads=256, blocks_per_cu=0, steps_per_sync=32, refill_min=64, queues=8, order=1, coop_steps=0),
    dict(block_threads=192, blocks_per_cu=2, steps_per_sync=7, refill_min=17, queues=3, order=1, coop_steps=0),
]


def _same(a, b):
    np.testing.assert_array_equal(a["status"], b["status"])
    np.testing.assert_array_equal(a["steps"], b["steps"])
    np.testing.assert_array_equal(a["iquv"], b["iquv"])


@pytest.mark.parametrize("storage,layout,precision", [("f32", "corner", PREC_FP32), ("bf16", "corner", PREC_MIXED),
                                                      ("f32", "soa", PREC_FP64)])
def test_tuning_does_not_change_the_image(renderer, storage, layout, precision):
    sc, br = scenes.scale_plain_scene(96, 80), scenes.scale_brick()
    db = renderer.upload(br, storage, layout)
    saved = renderer.get_tuning(precision)
    try:
        renderer.set_tuning(precision, coop_steps=0)
        base = renderer.render(sc, db, precision)
        for t in TUNINGS:
            renderer.set_tuning(precision, **t)
            _same(renderer.render(sc, db, precision), base)
    finally:
        renderer.set_tuning(precision, **saved)
        renderer.release(db)


@pytest.mark.parametrize("storage,precision,size,exact", [("f32", PREC_FP32, 1024, False), ("f32", PREC_FP32, 512, False),
                                                         ("bf16", PREC_MIXED, 1024, False), ("f32", PREC_FP32, 1024, True),
                                                         ("f16", PREC_MIXED, 512, True)])
def test_split_lone_rays_match_lone_rays(renderer, storage, precision, size, exact):
    """A lone ray whose geodesic and samples an idle wave takes over (coop_steps > 0) gives the image of the
    lone-wave path bit for bit: statuses, steps and IQUV equal on every pixel (both paths run the same step
    arithmetic, which rounds the same wherever it is inlined)."""
    sc, br = scenes.scale_plain_scene(size, size), scenes.scale_brick()
    sc = replace(sc, exact_step=exact)
    db = renderer.upload(br, storage, "corner")
    saved = renderer.get_tuning(precision)
    try:
        renderer.set_tuning(precision, coop_steps=0)
        base = renderer.render(sc, db, precision)
        assert renderer.last_coop() == (0, 0)
        for cs in (8, 32):
            renderer.set_tuning(precision, coop_steps=cs)
            out = renderer.render(sc, db, precision)
            split, timeouts = renderer.last_coop()
            n_diff = int((out["iquv"] != base["iquv"]).any(axis=-1).sum())
            print(f"{storage} {size} exact={exact}: coop_steps {cs}: {split} split rays, {timeouts} time-outs, "
                  f"{n_diff} pixels differ")
            assert split > 0 and timeouts == 0
            _same(out, base)
    finally:
        renderer.set_tuning(precision, **saved)
        renderer.release(db)


@pytest.mark.parametrize("block_threads", [64, 128])
def test_split_with_small_workgroups(renderer, block_threads):
    """With one wave per workgroup no helper exists: lone rays ask, nobody accepts, the request is withdrawn
    and the image is the lone-ray image; with two waves the split still happens."""
    sc, br = scenes.scale_plain_scene(512, 512), scenes.scale_brick()
    db = renderer.upload(br, "f32", "corner")
    saved = renderer.get_tuning(PREC_FP32)
    try:
        renderer.set_tuning(PREC_FP32, coop_steps=0)
        base = renderer.render(sc, db, PREC_FP32)
        renderer.set_tuning(PREC_FP32, block_threads=block_threads, coop_steps=8)
        out = renderer.render(sc, db, PREC_FP32)
        split, timeouts = renderer.last_coop()
        print(f"block_threads {block_threads}: {split} split rays, {timeouts} time-outs")
        assert timeouts == 0
        assert split == 0 if block_threads == 64 else split > 0
        _same(out, base)
    finally:
        renderer.set_tuning(PREC_FP32, **saved)
        renderer.release(db)


def test_autotuned_batch_gives_the_same_images(renderer):
    from bhgrrt import autotune

    bricks = scenes.batch_bricks()[:2]
    dbs = [renderer.upload(b, "bf16", "corner") for b in bricks]
    jobs = [(sc.with_size(w, w), dbs[k % 2]) for k, (sc, _), w in
            zip(range(4), scenes.batch_scenes(4, seed=11), (17, 40, 9, 31))]
    saved = renderer.get_tuning(PREC_MIXED)
    batch = Batch(renderer, jobs, PREC_MIXED)
    try:
        batch.launch(graph=False)
        before = batch.results()
        res = autotune.autotune(renderer, PREC_MIXED, lambda: batch.launch(graph=False),
                                space={"block_threads": (64, 256), "steps_per_sync": (1, 16), "blocks_per_cu": (0,)},
                                reps=1)
        assert res["best"] in [t for t, _ in res["trials"]] and len(res["trials"]) == 4
        assert renderer.get_tuning(PREC_MIXED)["steps_per_sync"] == res["best"]["steps_per_sync"]
        for graph in (False, True, True):
            batch.launch(graph=graph)
            for a, b in zip(batch.results(), before):
                _same(a, b)
    finally:
        batch.close()
        renderer.set_tuning(PREC_MIXED, **saved)
        for d in dbs:
            renderer.release(d)


def _check_fp64_against_oracle(renderer, sc, br, label):
    db = renderer.upload(br, "f32", "soa")
    try:
        out = renderer.render(sc, db, PREC_FP64)
    finally:
        renderer.release(db)
    ref = orender.render_rays(sc, br)
    st, steps = out["status"].ravel(), out["steps"].ravel()
    I = out["iquv"].reshape(-1, 4)
    ok = ((st | ref["status"]) & orender.ST_POLE_CLAMPED) == 0
    same = (st == ref["status"]) & (steps == ref["steps"])
    print(f"{label}: {sc.width}x{sc.height}, unflagged {ok.mean():.3f}, status+steps agreement {same[ok].mean():.4f}, "
          f"statuses {np.bincount(st & 7, minlength=5).tolist()}")
    assert same[ok].mean() >= 0.98 if ok.sum() >= 100 else same[ok].all()
    good = ok & same
    rel = np.abs(I[good] - ref["iquv"][good]) / np.maximum(np.abs(ref["iquv"][good, :1]), 1e-300)
    if good.any():
        print(f"{label}: rel I median {np.median(rel[:, 0]):.2e} p95 {np.quantile(rel[:, 0], 0.95):.2e} "
              f"max {rel[:, 0].max():.2e}")
        assert np.quantile(rel[:, 0], 0.95) <= 1e-12 and rel[:, 0].max() <= 1e-9
        assert np.max(np.abs(I[good, 1:] - ref["iquv"][good, 1:])) <= 1e-9 * max(np.abs(ref["iquv"][good, 0]).max(), 1.0)
    return out


@pytest.mark.parametrize("w,h", [(1, 1), (1, 9), (9, 1), (13, 7), (8, 8), (17, 16)])
def test_ragged_image_sizes_fp64_match_oracle(renderer, w, h):
    _check_fp64_against_oracle(renderer, scenes.scale_plain_scene(w, h), scenes.scale_brick(), f"ragged {w}x{h}")


def _scene(spin, inc, phi=0.0, size=20):
    cam = scenes.Camera.spanning(r=1000.0, inclination_degrees=inc, half_width_m=20.0, width=size, height=size,
                                 phi_degrees=phi)
    return scenes.Scene(spin=spin, camera=cam, width=size, height=size, step_mode=scenes.STEP_ADAPTIVE, step=0.05,
                        max_steps=4096, escape_radius=220.0)


@pytest.mark.parametrize("spin,inc,phi", [(0.0, 60.0, 0.0), (0.998, 17.0, 0.0), (0.9, 2.0, 0.0), (0.9, 89.0, 0.0),
                                          (0.5, 45.0, 73.0), (-0.7, 30.0, 0.0)])
def test_extreme_scenes_fp64_match_oracle(renderer, spin, inc, phi):
    br = scenes.batch_bricks()[0]       # 64x32x64 grid starting at 1.05 r_+ of a = 0.998: fits every spin
    _check_fp64_against_oracle(renderer, _scene(spin, inc, phi), br, f"a={spin} i={inc} phi={phi}")


def test_extreme_scenes_fp32_and_mixed_agree_with_fp64(renderer):
    """FP32 and mixed paths over the same extreme scenes: statuses agree on >= 97 % of unflagged rays."""
    br = scenes.batch_bricks()[0]
    d32, d16 = renderer.upload(br, "f32", "soa"), renderer.upload(br, "bf16", "corner")
    try:
        for spin, inc in ((0.0, 60.0), (0.998, 17.0), (0.9, 89.0), (-0.7, 30.0)):
            sc = _scene(spin, inc, size=32)
            o64 = renderer.render(sc, d32, PREC_FP64)
            for db, prec in ((d32, PREC_FP32), (d16, PREC_MIXED)):
                o = renderer.render(sc, db, prec)
                ok = ((o["status"] | o64["status"]) & orender.ST_POLE_CLAMPED) == 0
                agree = np.mean((o["status"] == o64["status"])[ok])
                print(f"a={spin} i={inc} prec={prec}: status agreement {agree:.4f}")
                assert agree >= 0.97
                assert np.all(np.isfinite(o["iquv"]))
    finally:
        renderer.release(d32)
        renderer.release(d16)


def test_invalid_arguments_are_errors(renderer):
    sc, br = scenes.scale_plain_scene(8, 8), scenes.scale_brick()
    db = renderer.upload(br, "f32", "corner")
    gone = renderer.upload(br, "f32", "soa")
    renderer.release(gone)
    try:
        with pytest.raises(RuntimeError, match="unknown brick id"):
            renderer.render(sc, gone, PREC_FP32)
        with pytest.raises(ValueError, match="precision must be"):
            renderer.render(sc, db, 5)
        with pytest.raises(ValueError, match="FP64 render needs a float32 brick in the SoA layout"):
            renderer.render(sc, db, PREC_FP64)
        with pytest.raises(ValueError, match="width and height must be positive"):
            renderer.render(sc.with_size(0, 8), db, PREC_FP32)
        with pytest.raises(ValueError, match="Kerr spin"):
            renderer.render(replace(sc, spin=1.0), db, PREC_FP32)
        inside = replace(sc, camera=scenes.Camera.from_degrees(r=1.2, inclination_degrees=60.0))
        with pytest.raises(ValueError, match="outside the outer horizon"):
            renderer.render(inside, db, PREC_FP32)
        with pytest.raises(ValueError, match="max_steps"):
            renderer.render(replace(sc, max_steps=0), db, PREC_FP32)
        with pytest.raises(ValueError, match="step must be finite"):
            renderer.render(replace(sc, step=math.nan), db, PREC_FP32)
        saved = renderer.get_tuning(PREC_FP32)
        for bad in (dict(block_threads=96), dict(queues=9), dict(refill_min=65), dict(order=2),
                    dict(steps_per_sync=5000)):
            with pytest.raises(ValueError):
                renderer.set_tuning(PREC_FP32, **bad)
            assert renderer.get_tuning(PREC_FP32) == saved
        other = renderer.upload(br, "bf16", "corner")
        try:
            with pytest.raises(ValueError, match="share storage and layout"):
                Batch(renderer, [(sc, db), (sc, other)], PREC_FP32)
        finally:
            renderer.release(other)
        # the context still renders after every rejected call
        out = renderer.render(sc, db, PREC_FP32)
        assert np.all(np.isfinite(out["iquv"]))
    finally:
        renderer.release(db)


def test_step_cap_matches_oracle(renderer):
    """max_steps = 1, 2, 7 with the camera inside the brick (r = 30), so the first steps already sample:
    every ray stops at the cap unless an earlier rule ends it, exactly as in the oracle."""
    br = scenes.scale_brick()
    db = renderer.upload(br, "f32", "soa")
    cam = scenes.Camera.spanning(r=30.0, inclination_degrees=60.0, half_width_m=12.0, width=24, height=16)
    try:
        for cap in (1, 2, 7):
            sc = scenes.Scene(spin=0.94, camera=cam, width=24, height=16, step_mode=scenes.STEP_ADAPTIVE, step=0.05,
                              max_steps=cap, escape_radius=220.0)
            out = renderer.render(sc, db, PREC_FP64)
            ref = orender.render_rays(sc, br)
            np.testing.assert_array_equal(out["status"].ravel(), ref["status"])
            np.testing.assert_array_equal(out["steps"].ravel(), ref["steps"])
            assert out["steps"].max() <= cap
            I, I0 = out["iquv"].reshape(-1, 4)[:, 0], ref["iquv"][:, 0]
            assert np.allclose(I, I0, rtol=1e-12, atol=1e-300)
    finally:
        renderer.release(db)


def test_large_image_fp32_against_fp64(renderer):
    """2048 x 2048 (4 M rays, 16x the 512^2 grid of queue slots): FP32 against FP64 over the whole image."""
    sc, br = scenes.scale_plain_scene(2048, 2048), scenes.scale_brick()
    d32, dc = renderer.upload(br, "f32", "soa"), renderer.upload(br, "f32", "corner")
    try:
        o64 = renderer.render(sc, d32, PREC_FP64)
        o32 = renderer.render(sc, dc, PREC_FP32)
    finally:
        renderer.release(d32)
        renderer.release(dc)
    ok = ((o64["status"] | o32["status"]) & orender.ST_POLE_CLAMPED) == 0
    assert np.mean((o64["status"] == o32["status"])[ok]) >= 0.99
    I32, I64 = o32["iquv"][..., 0][ok].astype(np.float64), o64["iquv"][..., 0][ok]
    e = float(np.abs(I32 - I64).sum() / np.abs(I64).sum())
    print(f"2048^2 fp32 vs fp64: rel L1 I {e:.2e}, flagged {1 - ok.mean():.4f}, longest ray {o32['steps'].max()}")
    assert e <= 1e-4
    assert o32["steps"].max() <= sc.max_steps and o32["steps"].min() >= 1


def test_image_size_limit_is_an_error(renderer):
    sc = scenes.scale_plain_scene(8, 8)
    db = renderer.upload(scenes.scale_brick(), "f32", "corner")
    try:
        with pytest.raises(ValueError, match="image is too large"):
            renderer.render(sc.with_size(1 << 16, 1 << 15), db, PREC_FP32)
    finally:
        renderer.release(db)
