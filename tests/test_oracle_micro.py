"""Pin the FP64 micro-kernel oracle (oracle/micro.py) to the reference's known answers and closed forms."""

import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import micro


def linear_fixture(nphi=2, step=0.125):
    """test_native_sampler_parity.py:22-35 (step 0.125) / lib.rs:384-397 (step 0.001)."""
    r_grid = np.array([0.0, 1.0])
    theta_grid = np.array([0.0, 1.0])
    phi_grid = np.array([0.0, 1.0]) if nphi == 2 else np.linspace(0.0, 2.0 * math.pi, nphi, endpoint=False)
    c = np.empty((2, 2, nphi, 11))
    for ir in range(2):
        for it in range(2):
            for ip in range(nphi):
                c[ir, it, ip] = 100.0 * ir + 10.0 * it + ip + np.arange(11) * step
    return c, r_grid, theta_grid, phi_grid


def sample_points():
    """test_native_sampler_parity.py:38-47."""
    return np.array([[0.25, 0.50, 0.75], [1.00, 1.00, 0.00], [-0.10, 0.50, 0.75], [0.25, 1.10, 0.75]])


def initial(n):
    out = np.zeros((n, 4))
    out[:, 0] = 1.0e-2
    out[:, 1] = 1.0e-3
    return out


# --- known answers from the reference's own tests -------------------------------

def test_lib_rs_scalar_absorption_kat():
    # lib.rs:407-414: j_I = 1, alpha_I = 0.5, ds = 0.1 -> I = 0.0975
    out = micro.stokes_rk2_brick(np.array([1.0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0, 0]), 0.1, np.zeros(4))
    assert abs(out[0] - 0.0975) < 1e-12
    assert out[1] == 0.0 and out[2] == 0.0 and out[3] == 0.0


def test_lib_rs_rejects_bad_shape():
    # lib.rs:417-419; test_native_stokes_parity.py:49-54
    with pytest.raises(ValueError, match="shape"):
        micro.stokes_rk2_brick(np.zeros(10), 0.1, np.zeros(4))
    with pytest.raises(ValueError, match="initial"):
        micro.stokes_rk2_brick(np.zeros((2, 2, 2, 11)), 0.05, np.zeros((2, 4)))


def test_lib_rs_center_kat():
    # lib.rs:421-434: centre of the 2x2x2 fixture (step 0.001) -> 55.5 and 55.51
    c, rg, tg, pg = linear_fixture(step=0.001)
    out = micro.sample_brick_trilinear(c, rg, tg, pg, [[0.5, 0.5, 0.5]])
    assert abs(out[0, 0] - 55.5) < 1e-12
    assert abs(out[0, 10] - 55.51) < 1e-12


def test_lib_rs_nan_outside():
    # lib.rs:436-448 and 450-464
    c, rg, tg, pg = linear_fixture(step=0.001)
    out = micro.sample_brick_trilinear(c, rg, tg, pg, [[-0.1, 0.5, 0.5], [0.5, 1.1, 0.5]])
    assert np.all(np.isnan(out))
    s = micro.sample_and_step_stokes(c, rg, tg, pg, [[-0.1, 0.5, 0.5]], 0.1, [1.0, 2.0, 3.0, 4.0])
    assert np.all(np.isnan(s))


def test_sampler_interpolates_in_bounds():
    # test_native_sampler_parity.py:57-65: (0.25, 0.5, 0.75) -> 30.75 + 0.125 k
    c, rg, tg, pg = linear_fixture()
    out = micro.sample_brick_trilinear(c, rg, tg, pg, [[0.25, 0.5, 0.75]])
    np.testing.assert_allclose(out[0], 30.75 + np.arange(11) * 0.125, rtol=1e-12, atol=1e-12)


def test_sampler_nan_outside_r_theta():
    # test_native_sampler_parity.py:68-83
    c, rg, tg, pg = linear_fixture()
    pts = [[-0.1, 0.5, 0.5], [1.1, 0.5, 0.5], [0.5, -0.1, 0.5], [0.5, 1.1, 0.5]]
    assert np.all(np.isnan(micro.sample_brick_trilinear(c, rg, tg, pg, pts)))


def test_valid_mask_kat():
    # test_native_sampler_parity.py:86-100
    _, rg, tg, _ = linear_fixture()
    pts = [[0.5, 0.5, 0.5], [-0.1, 0.5, 0.5], [0.5, 1.1, 0.5], [0.5, 0.5, 100.0]]
    np.testing.assert_array_equal(micro.sample_brick_valid_mask(rg, tg, pts), [True, False, False, True])


def test_periodic_phi_wrap():
    # test_native_sampler_parity.py:103-111
    c, rg, tg, pg = linear_fixture(nphi=4)
    base = micro.sample_brick_trilinear(c, rg, tg, pg, [[0.5, 0.5, math.pi / 4.0]])
    wrapped = micro.sample_brick_trilinear(c, rg, tg, pg, [[0.5, 0.5, 2.0 * math.pi + math.pi / 4.0]])
    np.testing.assert_allclose(wrapped, base, rtol=1e-12, atol=1e-12)


def test_sample_and_step_matches_sampler_then_rk2():
    # test_native_sampler_parity.py:156-167
    c, rg, tg, pg = linear_fixture()
    pts = sample_points()
    init = initial(4)
    sampled = micro.sample_brick_trilinear(c, rg, tg, pg, pts)
    valid = np.all(np.isfinite(sampled), axis=1)
    expected = np.full((4, 4), np.nan)
    expected[valid] = micro.stokes_rk2_brick(sampled[valid], 0.05, init[valid])
    out = micro.sample_and_step_stokes(c, rg, tg, pg, pts, 0.05, init)
    np.testing.assert_allclose(out, expected, rtol=1e-12, atol=1e-12, equal_nan=True)


def test_invalid_policies():
    # test_native_sampler_parity.py:170-212
    c, rg, tg, pg = linear_fixture()
    pts = sample_points()
    init = initial(4)
    valid = micro.sample_brick_valid_mask(rg, tg, pts)
    keep = micro.sample_and_step_stokes(c, rg, tg, pg, pts, 0.05, init, invalid_policy="initial")
    zero = micro.sample_and_step_stokes(c, rg, tg, pg, pts, 0.05, init, invalid_policy="zero")
    np.testing.assert_array_equal(keep[~valid], init[~valid])
    np.testing.assert_array_equal(zero[~valid], 0.0)
    with pytest.raises(ValueError, match="invalid_policy"):
        micro.sample_and_step_stokes(c, rg, tg, pg, pts, 0.05, invalid_policy="break")


def test_grid_endpoint_gives_unit_weight():
    # x == grid[-1] -> (n-2, n-1, 1.0) (native_kernels.py:140-141)
    g = np.array([0.0, 0.3, 1.0])
    assert micro.bracket_linear(g, 1.0) == (1, 2, 1.0)
    assert micro.bracket_linear(g, 0.0) == (0, 1, 0.0)
    assert micro.bracket_linear(g, 1.0 + 1e-16 * 0) is not None
    assert micro.bracket_linear(g, np.nextafter(1.0, 2.0)) is None


# --- closed forms / independent checks ------------------------------------------

def test_trilinear_reproduces_affine_data_exactly():
    rng = np.random.default_rng(1)
    rg = np.sort(rng.uniform(0, 5, 7)); tg = np.sort(rng.uniform(0, 3, 5))
    pg = np.linspace(0, 2 * math.pi, 6, endpoint=False)
    R, T = np.meshgrid(rg, tg, indexing="ij")
    c = np.empty((7, 5, 6, 11))
    for k in range(11):
        c[..., k] = ((1 + k) * R + (2 - k) * T)[..., None]   # affine in r, theta; constant in phi
    pts = np.column_stack([rng.uniform(rg[0], rg[-1], 200), rng.uniform(tg[0], tg[-1], 200), rng.uniform(-20, 20, 200)])
    out = micro.sample_brick_trilinear(c, rg, tg, pg, pts)
    expect = np.stack([(1 + k) * pts[:, 0] + (2 - k) * pts[:, 1] for k in range(11)], axis=1)
    np.testing.assert_allclose(out, expect, rtol=1e-12, atol=1e-12)


def test_phi_seam_interpolates_between_last_and_first():
    # periodic seam cell (native_kernels.py:157-161): phi between grid[-1] and grid[0] + 2 pi
    rg = np.array([0.0, 1.0]); tg = np.array([0.0, 1.0]); pg = np.linspace(0, 2 * math.pi, 4, endpoint=False)
    c = np.zeros((2, 2, 4, 11)); c[..., 3, :] = 8.0; c[..., 0, :] = 4.0
    phi = pg[-1] + 0.25 * (2 * math.pi - pg[-1])
    out = micro.sample_brick_trilinear(c, rg, tg, pg, [[0.5, 0.5, phi]])
    np.testing.assert_allclose(out[0], 0.75 * 8.0 + 0.25 * 4.0, rtol=1e-12)


def test_vectorised_sampler_equals_scalar():
    rng = np.random.default_rng(7)
    rg = np.geomspace(1.5, 40, 9); tg = np.linspace(0, math.pi, 7); pg = np.linspace(0, 2 * math.pi, 8, endpoint=False)
    c = rng.normal(size=(9, 7, 8, 11))
    pts = np.column_stack([rng.uniform(1.0, 45, 3000), rng.uniform(-0.3, 3.5, 3000), rng.uniform(-30, 30, 3000)])
    pts[:20, 0] = rg[rng.integers(0, 9, 20)]      # exact grid nodes
    pts[20:40, 2] = pg[rng.integers(0, 8, 20)] + 2 * math.pi * rng.integers(-3, 3, 20)
    a = micro.sample_brick_trilinear(c, rg, tg, pg, pts)
    b = micro.sample_brick_trilinear_vec(c, rg, tg, pg, pts)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    np.testing.assert_array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
    init = rng.normal(size=(3000, 4))
    for pol in ("nan", "initial", "zero"):
        x = micro.sample_and_step_stokes(c, rg, tg, pg, pts, 0.3, init, invalid_policy=pol)
        y = micro.sample_and_step_stokes_vec(c, rg, tg, pg, pts, 0.3, init, invalid_policy=pol)
        np.testing.assert_array_equal(x, y)


def _propagation_matrix(c):
    """K of synchrotron.py:82-94, written in the reference's matrix layout."""
    ai, aq, au, av, rv, rq, ru = c[4], c[5], c[6], c[7], c[8], c[9], c[10]
    return np.array([[ai, aq, au, av], [aq, ai, rv, -ru], [au, -rv, ai, rq], [av, ru, -rq, ai]])


def test_rk2_closed_form_and_convergence_to_matrix_exponential():
    rng = np.random.default_rng(3)
    c = np.concatenate([rng.uniform(0.1, 1.0, 4), rng.uniform(0.2, 0.6, 1), rng.uniform(-0.1, 0.1, 3),
                        rng.uniform(-0.8, 0.8, 3)])
    S0 = rng.normal(size=4)
    K = _propagation_matrix(c)
    j = c[:4]
    ds = 0.07
    # one step: S1 = S0 + ds (j - K S0) - ds^2/2 K (j - K S0)
    f0 = j - K @ S0
    np.testing.assert_allclose(micro.stokes_rk2_brick(c, ds, S0), S0 + ds * f0 - 0.5 * ds * ds * K @ f0, rtol=1e-13,
                               atol=1e-14)
    # many small steps converge to the exact step of polarized_transfer.py:51-70 (augmented expm)
    A = np.zeros((5, 5)); A[:4, :4] = -K; A[:4, 4] = j
    L = 1.3
    exact = (expm(A * L) @ np.append(S0, 1.0))[:4]
    errs = []
    for n in (50, 100):
        S = S0.copy()
        for _ in range(n):
            S = micro.stokes_rk2_brick(c, L / n, S)
        errs.append(np.max(np.abs(S - exact)))
    assert errs[1] < 1e-4
    assert 3.5 < errs[0] / errs[1] < 4.5   # second order


def test_faraday_rotation_kat():
    # test_synchrotron_polarized_transfer.py:19-24 (rho_V = 2 rotates Q into U, |Q,U| preserved)
    c = np.zeros(11); c[8] = 2.0
    S = np.array([0.0, 1.0, 0.0, 0.0])
    for _ in range(1000):
        S = micro.stokes_rk2_brick(c, 1e-4, S)
    assert abs(S[0]) < 1e-12
    assert abs(S[2]) > 0.05
    assert abs(np.hypot(S[1], S[2]) - 1.0) < 1e-9


def test_exact_step_closed_forms_and_rk2_limit():
    """oracle.micro.stokes_exact_step (polarized_transfer.py:51-70) against what the mathematics fixes:
    pure emission + absorption, pure Faraday rotation, the limit of many RK2 steps, and composition."""
    # emission + absorption: I(s) = I0 e^{-a s} + j/a (1 - e^{-a s})
    c = np.zeros((3, 11)); c[:, 0] = [1.0, 0.3, 2.0]; c[:, 4] = [0.5, 2.0, 1e-3]
    S0 = np.array([[0.2, 0, 0, 0], [0.0, 0, 0, 0], [5.0, 0, 0, 0]])
    ds = np.array([0.7, 3.0, 10.0])
    got = micro.stokes_exact_step(S0, c, ds)
    a, j = c[:, 4], c[:, 0]
    want = S0[:, 0] * np.exp(-a * ds) + j / a * (1 - np.exp(-a * ds))
    np.testing.assert_allclose(got[:, 0], want, rtol=1e-13)
    assert np.all(got[:, 1:] == 0.0)
    # Faraday rotation by rho_V: (Q, U) turns by rho_V s, V and I untouched
    c = np.zeros((1, 11)); c[0, 8] = 2.0
    s = np.array([0.3])
    got = micro.stokes_exact_step(np.array([[1.0, 1.0, 0.0, 0.25]]), c, s)[0]
    np.testing.assert_allclose(got, [1.0, math.cos(0.6), math.sin(0.6), 0.25], rtol=1e-13, atol=1e-15)
    # full coefficient sets: many RK2 steps converge to one exact step at second order
    rng = np.random.default_rng(11)
    c = np.concatenate([rng.uniform(0.1, 1.0, 4), rng.uniform(0.2, 0.6, 1), rng.uniform(-0.1, 0.1, 3),
                        rng.uniform(-0.8, 0.8, 3)])[None, :]
    S0 = rng.normal(size=(1, 4))
    exact = micro.stokes_exact_step(S0, c, np.array([1.1]))
    errs = []
    for n in (64, 128):
        S = S0.copy()
        for _ in range(n):
            S = micro.stokes_rk2_brick(c[0], 1.1 / n, S[0])[None, :]
        errs.append(np.max(np.abs(S - exact)))
    assert 3.5 < errs[0] / errs[1] < 4.5
    # composition: two half steps equal one step; ds <= 0 keeps S
    half = micro.stokes_exact_step(micro.stokes_exact_step(S0, c, np.array([0.55])), c, np.array([0.55]))
    np.testing.assert_allclose(half, exact, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(micro.stokes_exact_step(S0, c, np.array([0.0])), S0)
    np.testing.assert_array_equal(micro.stokes_exact_step(S0, c, np.array([-1.0])), S0)
