"""GPU side of the accelerated_renderer.py drop-in (bhgrrt.accelerated).

* Against the reference's own renderer (tests/golden/accelerated_vectors.npz, written by
  render_tile_stokes_bricks and trace_kerr_null_geodesic of the reference): the AcceleratedRenderConfig
  defaults (160x90, r=55, i=65 deg, fov 32 deg, RK4 step 0.06, 5000 steps; with the RK2 and the exact
  Stokes step) and the accelerated_cli.py defaults (64x36, step 0.08, 700 steps), a=0.85.
  FP64 (RK4, float32 sampling as the reference's NumPy sampler): statuses and step counts equal on every
  pixel (a ray ended by the optical-depth cap only up to that cap), Stokes IQUV within 1e-10 of |I| on
  every pixel not flagged at a pole and within 1e-6 on the flagged ones (measured: 2e-13 and 4e-11).
  FP32 (RK4, corner f32 brick), on pixels flagged in neither render: statuses on >= 97 %, relative L1
  error of I over each tile <= 5e-4, per-pixel relative I error <= 1.5e-3 at the 95th percentile,
  |dQ|, |dU|, |dV| <= 3e-3 of the tile's largest |I| (measured at most 1.1e-4, 3.2e-4 and 9e-4: FP32
  rounding accumulated over 900-4700 fixed RK4 steps).
* render_progressive_stokes_bricks: one batched launch; every frame has its level's size and equals the
  single render of that size bit for bit.
"""

import math
from pathlib import Path

import numpy as np
import pytest

from bhgrrt import accelerated as acc
from bhgrrt import scenes
from bhgrrt.render import PREC_FP32, PREC_FP64
from oracle import render as orender

pytestmark = pytest.mark.gpu

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "accelerated_vectors.npz")
CASES = ("default_a", "default_b", "exact", "cli")


def _config(name) -> acc.AcceleratedRenderConfig:
    """The reference config a fixture case was made with, built from this package's own defaults."""
    if name == "cli":   # accelerated_cli.py:26-33
        cfg = acc.AcceleratedRenderConfig(width=64, height=36, camera=scenes.Camera.from_degrees(r=55.0,
                                          inclination_degrees=65.0), step=0.08, max_steps=700)
    else:
        cfg = acc.AcceleratedRenderConfig(transfer=acc.PolarizedTransferConfig(use_exact_matrix_step=name == "exact"))
    w, h, cr, cth, cph, fov, roll, step, max_steps, esc, exact, tau_max, L = GOLD[f"{name}_config"]
    assert (cfg.width, cfg.height, cfg.step, cfg.max_steps, cfg.escape_radius) == (w, h, step, max_steps, esc)
    assert np.allclose([cfg.camera.r, cfg.camera.theta, cfg.camera.phi, cfg.camera.fov_y_degrees],
                       [cr, cth, cph, fov], rtol=1e-15, atol=0)
    assert cfg.transfer.use_exact_matrix_step == bool(exact) and cfg.integrator == "rk4"
    return cfg


def _brick():
    return scenes.Brick(GOLD["brick_r"], GOLD["brick_theta"], GOLD["brick_phi"], GOLD["brick_coeffs"])


def _tile(name, a):
    x0, y0, x1, y1 = (int(v) for v in GOLD[f"{name}_tile"])
    return a[y0:y1, x0:x1]


@pytest.fixture(scope="module")
def fixture_renders(renderer):
    br = _brick()
    d64, d32 = renderer.upload(br, "f32", "soa"), renderer.upload(br, "f32", "corner")
    out = {}
    try:
        for name in CASES:
            sc = acc.scene_from_config(_config(name), float(GOLD["spin"]))
            print(f"accelerated fixture {name}: {sc.width}x{sc.height} rk4", flush=True)
            out[name] = (renderer.render(sc, d64, PREC_FP64), renderer.render(sc, d32, PREC_FP32))
    finally:
        renderer.release(d64)
        renderer.release(d32)
    return out


@pytest.mark.parametrize("name", CASES)
def test_fp64_rk4_matches_reference_renderer(fixture_renders, name):
    o64 = fixture_renders[name][0]
    S, st, steps = _tile(name, o64["iquv"]), _tile(name, o64["status"]), _tile(name, o64["steps"])
    ref_S, ref_st, ref_steps = GOLD[f"{name}_stokes"], GOLD[f"{name}_status"], GOLD[f"{name}_steps"]
    code, flagged = st & 7, (st & orender.ST_POLE_CLAMPED) != 0
    opaque = code == orender.ST_OPAQUE
    np.testing.assert_array_equal(code[~opaque], ref_st[~opaque])
    np.testing.assert_array_equal(steps[~opaque], ref_steps[~opaque])
    assert np.all(steps[opaque] <= ref_steps[opaque])
    err = np.max(np.abs(S - ref_S), axis=-1) / np.abs(ref_S[..., 0])
    print(f"{name} fp64: max err/|I| unflagged {err[~flagged].max() if (~flagged).any() else 0:.2e}, "
          f"flagged {flagged.sum()} max {err[flagged].max() if flagged.any() else 0:.2e}")
    assert np.all(err[~flagged] <= 1e-10)
    assert np.all(err[flagged] <= 1e-6)


@pytest.mark.parametrize("name", CASES)
def test_fp32_rk4_against_reference_renderer(fixture_renders, name):
    """Pixels whose ray is flagged at a pole in the FP32 or the FP64 render are excluded (their steps through
    the theta clamp are not a geodesic in any precision) and counted."""
    o64, o32 = fixture_renders[name]
    S, st = _tile(name, o32["iquv"]).astype(np.float64), _tile(name, o32["status"])
    ok = ((st | _tile(name, o64["status"])) & orender.ST_POLE_CLAMPED) == 0
    ref_S, ref_st = GOLD[f"{name}_stokes"], GOLD[f"{name}_status"]
    same = (st & 7) == ref_st
    I, I0 = S[..., 0][ok], ref_S[..., 0][ok]
    rel_l1 = float(np.abs(I - I0).sum() / np.abs(I0).sum())
    pix = np.abs(I - I0) / np.abs(I0)
    quv = float(np.max(np.abs(S[..., 1:][ok] - ref_S[..., 1:][ok])) / np.abs(I0).max())
    print(f"{name} fp32: {ok.sum()} of {ok.size} pixels unflagged; statuses {same[ok].mean():.4f}, rel L1 I "
          f"{rel_l1:.2e}, pixel rel I p95 {np.quantile(pix, 0.95):.2e} max {pix.max():.2e}, "
          f"max |dQUV| / max|I| {quv:.2e}")
    assert ok.sum() >= 10
    assert same[ok].mean() >= 0.97
    assert rel_l1 <= 5e-4
    assert np.quantile(pix, 0.95) <= 1.5e-3
    assert quv <= 3e-3


def test_image_api_returns_the_fp64_render(renderer, fixture_renders):
    class Snapshot:
        spin_a = float(GOLD["spin"])

    img = acc.render_stokes_image_bricks(_config("default_b"), Snapshot(), _brick(), PREC_FP64, renderer=renderer)
    assert img.shape == (90, 160, 4) and img.dtype == np.float64
    np.testing.assert_array_equal(img, fixture_renders["default_b"][0]["iquv"])


def _c0_config(exact: bool) -> acc.AcceleratedRenderConfig:
    cam = scenes.Camera.spanning(r=1000.0, inclination_degrees=60.0, half_width_m=12.0, width=32, height=18)
    return acc.AcceleratedRenderConfig(width=32, height=18, camera=cam, step=scenes.SMALL_FIXED_STEP, max_steps=64,
                                       transfer=acc.PolarizedTransferConfig(use_exact_matrix_step=exact))


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
