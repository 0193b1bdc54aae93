"""Pin the fused render oracle (oracle/render.py).

* Against a scalar restatement with the reference's own structure: trace the
  whole geodesic first (kerr.py:337-384, with the RK2 step), then run the
  transfer loop over the stored states (accelerated_renderer.py:100-126).  The
  fused, vectorised, compacting oracle must give the same Stokes vectors.
* Against closed forms of the transfer: with alpha = 0 the RK2 Stokes update is
  exact, I = j * (sum of ds over in-brick segments); with constant absorption
  I -> (j / alpha)(1 - exp(-alpha s)); the optical-depth cap stops at tau >= 30.
"""

import math

import numpy as np

from bhgrrt import scenes
from oracle import kerr, micro, render


def _trace_then_transfer(scene, brick, px, py):
    a = scene.spin
    W, H = scene.width, scene.height
    ndc_x = 2.0 * (px + 0.5) / W - 1.0
    ndc_y = 1.0 - 2.0 * (py + 0.5) / H
    y, pt, pph = kerr.camera_ray_initial_state(scene.camera, a, ndc_x, ndc_y, W / H)
    rp = kerr.horizon_radius(a)
    states = [y.copy()]
    prev_r = y[1]
    for _ in range(scene.max_steps):
        h = scene.step * (y[1] - rp) if scene.step_mode == 1 else scene.step
        y = kerr.rk2_step(y, h, a, pt, pph)
        if not np.all(np.isfinite(y)):
            break
        states.append(y.copy())
        if y[1] <= rp * (1.0 + 2.0e-4):
            break
        if y[1] > scene.escape_radius and y[1] > prev_r:
            break
        prev_r = y[1]
    if not np.allclose(states[-1], y):
        states.append(y.copy())
    states = np.vstack(states)
    c64 = brick.coeffs.astype(np.float64)
    rg, tg, pg = (g.astype(np.float64) for g in (brick.r, brick.theta, brick.phi))
    S = np.zeros(4)
    tau = 0.0
    for i in range(1, states.shape[0]):
        mid = 0.5 * (states[i - 1] + states[i])
        if micro.bracket_linear(rg, float(mid[1])) is None or micro.bracket_linear(tg, float(mid[2])) is None:
            continue
        c = micro.sample_one(c64, rg, tg, pg, mid[1:4])
        dlam = float(np.linalg.norm(states[i][1:4] - states[i - 1][1:4]))
        ds = max(scene.affine_to_length_cm * max(dlam, 1.0e-14), 0.0)
        S = micro.stokes_rk2_brick(c, ds, S)
        tau += max(c[4], 0.0) * ds
        if tau >= scene.max_optical_depth:
            break
    return np.nan_to_num(S, nan=0.0, posinf=0.0, neginf=0.0)


def test_fused_oracle_matches_trace_then_transfer_small_scene():
    sc, br = scenes.small_plain_scene(), scenes.small_brick()
    out = render.render_image(sc, br)
    for py in range(sc.height):
        for px in range(sc.width):
            if out["status"][py, px] & render.ST_POLE_CLAMPED:
                continue
            ref = _trace_then_transfer(sc, br, px, py)
            np.testing.assert_allclose(out["iquv"][py, px], ref, rtol=1e-9, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_fused_oracle_matches_trace_then_transfer_adaptive_scene():
    sc = scenes.scale_plain_scene(48, 48)
    br = scenes.scale_brick(24, 12, 24)
    out = render.render_image(sc, br)
    rng = np.random.default_rng(5)
    picks = [(int(x), int(y)) for x, y in zip(rng.integers(0, 48, 60), rng.integers(0, 48, 60))]
    n = 0
    for px, py in picks:
        if out["status"][py, px] & render.ST_POLE_CLAMPED:
            continue
        ref = _trace_then_transfer(sc, br, px, py)
        np.testing.assert_allclose(out["iquv"][py, px], ref, rtol=1e-7, atol=1e-12)
        n += 1
    assert n > 40


def _uniform_brick(j, alpha, spin=0.94, n=(6, 5, 6)):
    rp = kerr.horizon_radius(spin)
    r = np.geomspace(1.05 * rp, 40.0, n[0])
    th = np.linspace(0.0, math.pi, n[1])
    ph = np.linspace(0.0, 2 * math.pi, n[2], endpoint=False)
    c = np.zeros(n + (11,), dtype=np.float32)
    c[..., 0] = j
    c[..., 4] = alpha
    return scenes.Brick(r, th, ph, c)


def test_transfer_closed_forms():
    sc = scenes.scale_plain_scene(24, 24)
    pure = render.render_image(sc, _uniform_brick(0.5, 0.0))
    ok = (pure["status"] & render.ST_POLE_CLAMPED) == 0
    s = pure["iquv"][..., 0] / 0.5          # path length through the brick (sum of coordinate chords)
    assert np.median(s[ok]) > 30.0         # rays cross the r <= 40 sphere (a few graze a pole and skip it)
    np.testing.assert_array_equal(pure["iquv"][..., 1:][ok], 0.0)
    alpha = 0.002
    absorbed = render.render_image(sc, _uniform_brick(0.5, alpha))
    expect = 0.5 / alpha * (1.0 - np.exp(-alpha * s))
    np.testing.assert_allclose(absorbed["iquv"][..., 0][ok], expect[ok], rtol=2e-4)


def test_optical_depth_cap_stops_the_ray():
    import dataclasses

    sc = dataclasses.replace(scenes.scale_plain_scene(16, 16), max_optical_depth=5.0)
    out = render.render_image(sc, _uniform_brick(1.0, 0.5))
    ok = (out["status"] & render.ST_POLE_CLAMPED) == 0
    opaque = (out["status"] & 7) == render.ST_OPAQUE
    assert np.mean(opaque[ok]) > 0.9
    # RK2's fixed point for constant coefficients is j / alpha; tau >= 5 leaves I within e^-5 of it
    np.testing.assert_allclose(out["iquv"][..., 0][ok & opaque], 2.0, rtol=2e-2)


def test_status_codes_and_steps_are_consistent():
    sc = scenes.scale_plain_scene(32, 32)
    out = render.render_image(sc, scenes.scale_brick(16, 8, 16))
    code = out["status"] & 7
    assert set(np.unique(code)) <= {1, 2, 4}
    assert np.all(out["steps"] >= 1) and np.all(out["steps"] <= sc.max_steps)
    assert 0 < np.sum(code == 1) < code.size
    # the shadow (captured rays) is the inside of the Bardeen critical curve: pixels more than 1.5 pixel
    # widths inside it are captured and pixels that far outside are not (the image's x is -alpha)
    al, be = kerr.critical_curve(sc.spin, sc.camera.theta, n=4000)
    px, py = np.meshgrid(np.arange(sc.width), np.arange(sc.height))
    ndc_x, ndc_y = scenes.pixel_ndc(sc, px, py)
    tan_y = math.tan(math.radians(sc.camera.fov_y_degrees) / 2)
    x_m, y_m = -sc.camera.r * ndc_x * tan_y * sc.aspect, sc.camera.r * ndc_y * tan_y
    inside = _inside_polygon(x_m, y_m, al, be)
    dist = np.min(np.hypot(x_m[..., None] - al, y_m[..., None] - be), axis=-1)
    far = dist > 1.5 * 2.0 * sc.camera.r * tan_y * sc.aspect / sc.width
    ok = (out["status"] & render.ST_POLE_CLAMPED) == 0
    assert np.sum(far & inside & ok) >= 10 and np.sum(far & ~inside & ok) > 200
    assert np.all((code == 1)[far & inside & ok]) and not np.any((code == 1)[far & ~inside & ok])


def _inside_polygon(x, y, px, py):
    """Even-odd ray casting of points (x, y) against the closed polygon (px, py)."""
    inside = np.zeros(np.shape(x), dtype=bool)
    qx, qy = np.roll(px, 1), np.roll(py, 1)
    for x0, y0, x1, y1 in zip(px, py, qx, qy):
        crosses = (y0 > y) != (y1 > y)
        xc = x0 + (y - y0) * (x1 - x0) / np.where(y1 != y0, y1 - y0, 1.0)
        inside ^= crosses & (x < xc)
    return inside
