"""GPU side of the accelerated_renderer.py drop-in (bhgrrt.accelerated).

* render_stokes_image_bricks(cfg, snapshot, bricks) in FP64 equals the FP64 oracle of the same scene
  (statuses and step counts exact, median relative I error <= 1e-10), with the RK2 and with the exact
  Stokes step.
* render_progressive_stokes_bricks: one batched launch; every frame has its level's size and equals the
  single render of that size bit for bit.
"""

import math

import numpy as np
import pytest

from bhgrrt import accelerated as acc
from bhgrrt import scenes
from bhgrrt.render import PREC_FP32, PREC_FP64
from oracle import render as orender

pytestmark = pytest.mark.gpu


def _c0_config(exact: bool) -> acc.AcceleratedRenderConfig:
    cam = scenes.Camera.spanning(r=1000.0, inclination_degrees=60.0, half_width_m=12.0, width=32, height=18)
    return acc.AcceleratedRenderConfig(width=32, height=18, camera=cam, step=scenes.SMALL_FIXED_STEP, max_steps=64,
                                       transfer=acc.PolarizedTransferConfig(use_exact_matrix_step=exact))


@pytest.mark.parametrize("exact", [False, True])
def test_image_api_matches_oracle(renderer, exact):
    cfg, br = _c0_config(exact), scenes.small_brick()

    class Snapshot:
        spin_a = 0.9

    img = acc.render_stokes_image_bricks(cfg, Snapshot(), br, PREC_FP64, renderer=renderer)
    assert img.shape == (18, 32, 4) and img.dtype == np.float64
    ref = orender.render_image(acc.scene_from_config(cfg, 0.9), br)
    I, I0 = img[..., 0], ref["iquv"][..., 0]
    rel = np.abs(I - I0) / np.maximum(np.abs(I0), 1e-300)
    print(f"image api exact={exact}: median rel {np.median(rel):.2e}")
    assert np.median(rel) <= 1e-10


def test_progressive_frames_equal_single_renders(renderer):
    br = scenes.scale_brick()
    cam = scenes.Camera.from_degrees(r=1000.0, inclination_degrees=17.0, fov_y_degrees=math.degrees(2 * math.atan(0.02)))
    cfg = acc.AcceleratedRenderConfig(width=96, height=54, camera=cam, step=0.05, step_mode=scenes.STEP_ADAPTIVE,
                                      max_steps=4096, progressive_min_width=24)
    frames = acc.render_progressive_stokes_bricks(cfg, 0.94, br, PREC_FP32, renderer=renderer)
    assert [(f.level, f.width, f.height) for f in frames] == [(0, 24, 13), (1, 48, 27), (2, 96, 54)]
    db = renderer.upload(br, "f32", "corner")
    saved = renderer.get_tuning(PREC_FP32)
    # the batched launch has no two-wave split of lone rays: compare with single renders without it
    renderer.set_tuning(PREC_FP32, coop_steps=0)
    try:
        for f in frames:
            single = renderer.render(acc.scene_from_config(cfg, 0.94, f.width, f.height), db, PREC_FP32)
            assert f.stokes.shape == (f.height, f.width, 4)
            np.testing.assert_array_equal(f.stokes, single["iquv"].astype(np.float64))
            assert np.abs(f.stokes[..., 0]).sum() > 0
    finally:
        renderer.set_tuning(PREC_FP32, **saved)
        renderer.release(db)


def test_float16_brick_file_renders_like_the_source_brick(renderer, tmp_path):
    """A float16 brick file in the reference's schema, loaded and uploaded as FP16, gives the same image
    as the source brick uploaded as FP16 (both round each value to the same half)."""
    src = scenes.scale_brick(64, 32, 64, seed=5)
    p = tmp_path / "half.npz"
    np.savez_compressed(p, r=src.r, theta=src.theta, phi=src.phi, coeffs=src.coeffs.astype(np.float16),
                        coeff_names=np.array(scenes.COEFF_NAMES), nu_hz=np.array(230e9), spin_a=np.array(0.94),
                        dtype_name=np.array("float16"))
    back = scenes.Brick.load_npz(p)
    sc = scenes.scale_plain_scene(40, 40)
    a, b = renderer.upload(src, "f16", "corner"), renderer.upload(back, "f16", "corner")
    try:
        oa, ob = renderer.render(sc, a, PREC_FP32), renderer.render(sc, b, PREC_FP32)
    finally:
        renderer.release(a)
        renderer.release(b)
    np.testing.assert_array_equal(oa["iquv"], ob["iquv"])
    np.testing.assert_array_equal(oa["status"], ob["status"])
    assert np.abs(oa["iquv"][..., 0]).sum() > 0


def test_tile_render_is_the_block_of_the_image(renderer):
    cfg, br = _c0_config(False), scenes.small_brick()
    img = acc.render_stokes_image_bricks(cfg, 0.9, br, PREC_FP64, renderer=renderer)
    tiles = acc.generate_tiles(cfg.width, cfg.height, tile_size=20)
    assert len(tiles) == 2
    for t in tiles:
        block = acc.render_tile_stokes_bricks(cfg, 0.9, br, t, PREC_FP64, renderer=renderer)
        np.testing.assert_array_equal(block, img[t.y0:t.y1, t.x0:t.x1])
