"""Host side of the accelerated_renderer.py drop-in (CPU): progressive levels and tiles against values
worked out by hand from tiles.py:30-50, and the config -> scene mapping."""

import math

from bhgrrt import accelerated as acc
from bhgrrt import scenes


def test_progressive_levels_worked_examples():
    # 1920 // 2 = 960 >= 480, 1920 // 4 = 480 >= 480, 1920 // 8 = 240 < 480: coarsest scale 4
    assert acc.progressive_levels(1920, 1080, 480) == [(0, 480, 270), (1, 960, 540), (2, 1920, 1080)]
    # 160 // 2 = 80 < 480: one level at full size
    assert acc.progressive_levels(160, 90, 480) == [(0, 160, 90)]
    # 1000 // 2 = 500 >= 480, 1000 // 4 = 250 < 480
    assert acc.progressive_levels(1000, 500, 480) == [(0, 500, 250), (1, 1000, 500)]
    # heights floor at 1
    assert acc.progressive_levels(4096, 3, 512) == [(0, 512, 1), (1, 1024, 1), (2, 2048, 1), (3, 4096, 3)]


def test_generate_tiles_cover_the_image_once():
    tiles = acc.generate_tiles(150, 70, tile_size=64)
    assert [(t.x0, t.y0, t.x1, t.y1) for t in tiles[:3]] == [(0, 0, 64, 64), (64, 0, 128, 64), (128, 0, 150, 64)]
    assert sum(t.width * t.height for t in tiles) == 150 * 70
    assert tiles[-1].width == 150 - 128 and tiles[-1].height == 70 - 64


def test_config_maps_to_the_scene_the_reference_renders():
    cfg = acc.AcceleratedRenderConfig()
    assert (cfg.width, cfg.height, cfg.step, cfg.max_steps, cfg.escape_radius) == (160, 90, 0.06, 5000, 220.0)
    assert cfg.transfer.use_exact_matrix_step is False and cfg.transfer.max_optical_depth == 30.0
    assert acc.PolarizedTransferConfig().use_exact_matrix_step is True
    assert math.isclose(cfg.camera.theta, math.radians(65.0)) and cfg.camera.r == 55.0
    assert cfg.camera.fov_y_degrees == 32.0

    class Snap:
        spin_a = 0.7

    sc = acc.scene_from_config(cfg, Snap())
    assert (sc.spin, sc.width, sc.height, sc.step_mode, sc.step) == (0.7, 160, 90, scenes.STEP_FIXED, 0.06)
    assert sc.affine_to_length_cm == 1.0 and sc.exact_step is False
    sc2 = acc.scene_from_config(acc.AcceleratedRenderConfig(
        transfer=acc.PolarizedTransferConfig(affine_to_length_cm=2.5, max_optical_depth=7.0)), 0.3, 40, 20)
    assert (sc2.spin, sc2.width, sc2.height, sc2.affine_to_length_cm, sc2.max_optical_depth, sc2.exact_step) == \
        (0.3, 40, 20, 2.5, 7.0, True)
