"""Pin the oracle's RK4 render against the reference's own CPU renderer.

tests/golden/accelerated_vectors.npz holds, for every pixel of four tiles, the
Stokes vector that python/blackhole_sim/accelerated_renderer.py:129-150
(render_tile_stokes_bricks: camera launch, trace_kerr_null_geodesic with rk4_step,
integrate_stokes_from_bricks_trace) returned, and the status and step count of
trace_kerr_null_geodesic (kerr.py:337-384) for the same ray; the cases are the
AcceleratedRenderConfig defaults (two tiles, and one with the exact Stokes step)
and the accelerated_cli.py defaults.  tests/golden/make_accelerated_vectors.py
wrote it.

The oracle restates the same algorithm (the matrix-form RHS evaluated per
component; the brick sampled in float32 as NumPy runs the reference's sampler on
a float32 brick), so it agrees to rounding: statuses and step counts exactly,
Stokes vectors to 1e-10 of |I| on every pixel, pole-flagged ones included (both
sides take the same clamped steps).  A ray the oracle ends at the optical-depth cap
(status 4) has a trace in the reference that runs on (its transfer stops
there), so only its Stokes vector is compared.
"""

import math
from pathlib import Path

import numpy as np
import pytest

from bhgrrt import scenes
from oracle import render

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "accelerated_vectors.npz")
CASES = ("default_a", "default_b", "exact", "cli")


def fixture_brick():
    return scenes.Brick(GOLD["brick_r"], GOLD["brick_theta"], GOLD["brick_phi"], GOLD["brick_coeffs"])


def fixture_scene(name):
    w, h, cr, cth, cph, fov, roll, step, max_steps, esc, exact, tau_max, L = GOLD[f"{name}_config"]
    cam = scenes.Camera(r=float(cr), theta=float(cth), phi=float(cph), fov_y_degrees=float(fov),
                        roll_degrees=float(roll))
    return scenes.Scene(spin=float(GOLD["spin"]), camera=cam, width=int(w), height=int(h),
                        step_mode=scenes.STEP_FIXED, step=float(step), max_steps=int(max_steps),
                        escape_radius=float(esc), affine_to_length_cm=float(L), max_optical_depth=float(tau_max),
                        exact_step=bool(exact), integrator="rk4",
                        numpy32_sampling=True)


def fixture_pixels(name, scene):
    x0, y0, x1, y1 = (int(v) for v in GOLD[f"{name}_tile"])
    ys, xs = np.mgrid[y0:y1, x0:x1]
    return (ys * scene.width + xs).ravel()


def compare(name, iquv, steps, status, tol):
    """Statuses and steps equal to the reference trace (but for rays ended by the optical-depth cap);
    every Stokes component within tol of the reference's |I| (or of tol itself where I is 0)."""
    ref_S = GOLD[f"{name}_stokes"].reshape(-1, 4)
    ref_status = GOLD[f"{name}_status"].ravel()
    ref_steps = GOLD[f"{name}_steps"].ravel()
    code = status & 7
    opaque = code == render.ST_OPAQUE
    np.testing.assert_array_equal(code[~opaque], ref_status[~opaque])
    np.testing.assert_array_equal(steps[~opaque], ref_steps[~opaque])
    assert np.all(steps[opaque] <= ref_steps[opaque])
    scale = np.maximum(np.abs(ref_S[:, :1]), 1e-300)
    err = np.abs(iquv - ref_S) / scale
    err[ref_S[:, :1].repeat(4, 1) == 0] = np.abs(iquv - ref_S)[ref_S[:, :1].repeat(4, 1) == 0]
    return float(err.max())


@pytest.mark.parametrize("name", CASES)
def test_oracle_rk4_render_matches_reference_renderer(name):
    sc, br = fixture_scene(name), fixture_brick()
    pix = fixture_pixels(name, sc)
    out = render.render_rays(sc, br, pix)
    worst = compare(name, out["iquv"], out["steps"], out["status"], 1e-10)
    print(name, "max rel err vs reference", worst)
    assert worst <= 1e-10


def test_fixture_covers_the_statuses_and_carries_signal():
    assert sum(int(np.sum(GOLD[f"{n}_status"] == 1)) for n in CASES) > 100   # captured
    assert sum(int(np.sum(GOLD[f"{n}_status"] == 2)) for n in CASES) > 100   # escaped
    assert int(np.sum(GOLD["cli_status"] == 0)) > 100                         # out of steps
    for n in CASES:
        assert np.all(np.abs(GOLD[f"{n}_stokes"][..., 0]) > 0)


def test_rk4_step_is_fourth_order_and_keeps_carter_constant():
    """kerr.py:306-313 restated: halving h cuts the end-point error by ~16 and the drift of the Hamiltonian
    and of Carter's constant Q by ~16; RK2 (the render's other rule) only by ~4."""
    from oracle import kerr

    a = 0.85
    cam = scenes.Camera.from_degrees(r=30.0, inclination_degrees=70.0, fov_y_degrees=30.0)
    y0, pt, pph = kerr.camera_ray_initial_state(cam, a, 0.3, 0.2, 1.0)
    Q0 = kerr.carter_constant_vec(y0[2], y0[5], pt, pph, a)
    lam = 16.0

    def run(step, n):
        y = y0.copy()
        for _ in range(n):
            y = step(y, lam / n, a, pt, pph)
        return y

    fine = run(kerr.rk4_step, 4096)
    e4 = [np.max(np.abs(run(kerr.rk4_step, n)[1:] - fine[1:])) for n in (64, 128)]
    assert 13.0 < e4[0] / e4[1] < 19.0, e4
    q4 = [abs(kerr.carter_constant_vec(y[2], y[5], pt, pph, a) - Q0) for y in (run(kerr.rk4_step, n) for n in (64, 128))]
    assert 13.0 < q4[0] / q4[1] < 19.0, q4
    e2 = [np.max(np.abs(run(kerr.rk2_step, n)[1:] - fine[1:])) for n in (128, 256)]
    assert 3.0 < e2[0] / e2[1] < 5.0, e2
    assert abs(kerr.hamiltonian_null(run(kerr.rk4_step, 128), a, pt, pph)) < 1e-7
