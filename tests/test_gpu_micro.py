"""GPU parity of the FP64 drop-in micro-kernels against the oracle.

Gate of the reference: rtol = atol = 1e-12 with equal_nan (native_kernels.py:12-15,
docs/PERFORMANCE_PLAN.md:30-31); the valid mask is bit-exact.  The cases are
the reference's own parity tests (test_native_sampler_parity.py,
test_native_stokes_parity.py, lib.rs:399-464) plus large seeded sets with
out-of-domain points, exact grid nodes and wrapped phi.
"""

import math

import numpy as np
import pytest

import blackhole_native
from bhgrrt import native_kernels as nk
from bhgrrt import scenes
from oracle import micro

pytestmark = pytest.mark.gpu
TOL = dict(rtol=1e-12, atol=1e-12, equal_nan=True)


def linear_fixture(nphi=2, step=0.125):
    r = np.array([0.0, 1.0]); t = np.array([0.0, 1.0])
    p = np.array([0.0, 1.0]) if nphi == 2 else np.linspace(0.0, 2 * math.pi, nphi, endpoint=False)
    c = np.empty((2, 2, nphi, 11))
    for ir in range(2):
        for it in range(2):
            for ip in range(nphi):
                c[ir, it, ip] = 100.0 * ir + 10.0 * it + ip + np.arange(11) * step
    return c, r, t, p


def initial(n):
    out = np.zeros((n, 4)); out[:, 0] = 1e-2; out[:, 1] = 1e-3
    return out


def test_lib_rs_kats_through_drop_in_module():
    out = blackhole_native.stokes_rk2_brick([1.0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0, 0], 0.1, [0.0] * 4)
    assert abs(out[0] - 0.0975) < 1e-12 and out[1] == 0.0 and out[2] == 0.0 and out[3] == 0.0
    c, r, t, p = linear_fixture(step=0.001)
    s = blackhole_native.sample_brick_trilinear(c.ravel().tolist(), [0, 1], [0, 1], [0, 1], [0.5, 0.5, 0.5])
    assert len(s) == 11 and abs(s[0] - 55.5) < 1e-12 and abs(s[10] - 55.51) < 1e-12
    s = blackhole_native.sample_brick_trilinear(c.ravel(), [0, 1], [0, 1], [0, 1], [-0.1, 0.5, 0.5, 0.5, 1.1, 0.5])
    assert len(s) == 22 and np.all(np.isnan(s))
    s = blackhole_native.sample_and_step_stokes(c.ravel(), [0, 1], [0, 1], [0, 1], [-0.1, 0.5, 0.5], 0.1, [1, 2, 3, 4])
    assert len(s) == 4 and np.all(np.isnan(s))


def test_native_stokes_matches_reference():
    # test_native_stokes_parity.py:31-46
    coeffs = nk.deterministic_stokes_coefficients(4, 3, 5)
    init = np.zeros(coeffs.shape[:-1] + (4,)); init[..., 0] = 1e-2; init[..., 1] = 1e-3
    np.testing.assert_allclose(nk.stokes_rk2_brick(coeffs, 0.05, init), micro.stokes_rk2_brick(coeffs, 0.05, init),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(nk.stokes_rk2_brick(coeffs, 0.05, [1.0, 2.0, 3.0, 4.0]),
                               micro.stokes_rk2_brick(coeffs, 0.05, np.array([1.0, 2.0, 3.0, 4.0])), rtol=1e-12,
                               atol=1e-12)


def test_native_sampler_matches_reference():
    # test_native_sampler_parity.py:130-153
    c, r, t, p = linear_fixture(nphi=4)
    pts = np.array([[0.25, 0.5, math.pi / 4], [0.5, 0.25, 2 * math.pi + math.pi / 4], [-0.1, 0.5, math.pi / 4],
                    [0.5, 1.1, math.pi / 4]])
    np.testing.assert_allclose(nk.sample_brick_trilinear(c, r, t, p, pts), micro.sample_brick_trilinear(c, r, t, p, pts), **TOL)
    one = nk.sample_brick_trilinear(*linear_fixture(), [[0.25, 0.5, 0.75]])
    np.testing.assert_allclose(one[0], 30.75 + np.arange(11) * 0.125, rtol=1e-12, atol=1e-12)


def test_native_sample_and_step_and_policies_match_reference():
    # test_native_sampler_parity.py:215-272
    c, r, t, p = linear_fixture(nphi=4)
    pts = np.array([[0.25, 0.5, math.pi / 4], [0.5, 0.25, 2 * math.pi + math.pi / 4], [-0.1, 0.5, math.pi / 4]])
    init = initial(3)
    np.testing.assert_allclose(nk.sample_and_step_stokes(c, r, t, p, pts, 0.05, init),
                               micro.sample_and_step_stokes(c, r, t, p, pts, 0.05, init), **TOL)
    pts2 = np.array([[0.25, 0.5, math.pi / 4], [-0.1, 0.5, math.pi / 4], [0.5, 1.1, math.pi / 4]])
    for pol in ("nan", "initial", "zero"):
        np.testing.assert_allclose(nk.sample_and_step_stokes(c, r, t, p, pts2, 0.05, init, invalid_policy=pol),
                                   micro.sample_and_step_stokes(c, r, t, p, pts2, 0.05, init, invalid_policy=pol), **TOL)


def test_valid_mask_kat_and_drop_in():
    _, r, t, _ = linear_fixture()
    pts = [[0.5, 0.5, 0.5], [-0.1, 0.5, 0.5], [0.5, 1.1, 0.5], [0.5, 0.5, 100.0]]
    np.testing.assert_array_equal(nk.sample_brick_valid_mask(r, t, pts), [True, False, False, True])
    np.testing.assert_array_equal(blackhole_native.sample_brick_valid_mask(r, t, np.ravel(pts)), [True, False, False, True])


def _random_case(seed, n, nr=9, nt=7, np_=8):
    rng = np.random.default_rng(seed)
    rg = np.geomspace(1.4, 40, nr); tg = np.linspace(0, math.pi, nt); pg = np.linspace(0, 2 * math.pi, np_, endpoint=False)
    c = rng.normal(size=(nr, nt, np_, 11))
    pts = np.column_stack([rng.uniform(1.0, 45, n), rng.uniform(-0.3, 3.5, n), rng.uniform(-40, 40, n)])
    k = n // 10
    pts[:k, 0] = rg[rng.integers(0, nr, k)]                 # exact r nodes, including the last
    pts[k:2 * k, 1] = tg[rng.integers(0, nt, k)]            # exact theta nodes
    pts[2 * k:3 * k, 2] = pg[rng.integers(0, np_, k)] + 2 * math.pi * rng.integers(-4, 4, k)
    pts[3 * k:3 * k + 8, 0] = [rg[0], rg[-1], np.nextafter(rg[0], 0), np.nextafter(rg[-1], 99)] * 2
    return c, rg, tg, pg, pts, rng


@pytest.mark.parametrize("seed,n", [(1, 20000), (2, 200000)])
def test_large_random_parity(seed, n):
    c, rg, tg, pg, pts, rng = _random_case(seed, n)
    np.testing.assert_array_equal(nk.sample_brick_valid_mask(rg, tg, pts), micro.sample_brick_valid_mask(rg, tg, pts))
    np.testing.assert_allclose(nk.sample_brick_trilinear(c, rg, tg, pg, pts),
                               micro.sample_brick_trilinear_vec(c, rg, tg, pg, pts), **TOL)
    init = rng.normal(size=(n, 4))
    for pol in ("nan", "initial", "zero"):
        np.testing.assert_allclose(nk.sample_and_step_stokes(c, rg, tg, pg, pts, 0.37, init, invalid_policy=pol),
                                   micro.sample_and_step_stokes_vec(c, rg, tg, pg, pts, 0.37, init, invalid_policy=pol),
                                   **TOL)
    cells = rng.normal(size=(n, 11))
    np.testing.assert_allclose(nk.stokes_rk2_brick(cells, 0.21, init), micro.stokes_rk2_brick(cells, 0.21, init),
                               rtol=1e-12, atol=1e-12)


def test_empty_inputs():
    c, r, t, p = linear_fixture()
    assert nk.sample_brick_trilinear(c, r, t, p, np.zeros((0, 3))).shape == (0, 11)
    assert nk.sample_and_step_stokes(c, r, t, p, np.zeros((0, 3)), 0.1).shape == (0, 4)
    assert nk.sample_brick_valid_mask(r, t, np.zeros((0, 3))).shape == (0,)
    assert len(blackhole_native.sample_brick_trilinear(c.ravel(), r, t, p, [])) == 0


def _device_sample(renderer, coeffs, grids, pts):
    """bh_sample_brick_trilinear_device on device copies of the arrays; returns the (n, 11) host result."""
    import ctypes as C

    from bhgrrt import _lib

    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (coeffs, *grids, pts)]
    bufs = [renderer.alloc(a.nbytes) for a in arrs]
    for b, a in zip(bufs, arrs):
        b.from_host(a)
    n = pts.shape[0]
    out = renderer.alloc(n * 11 * 8)
    try:
        _lib.check(renderer._lib.bh_sample_brick_trilinear_device(
            renderer.ctx, bufs[0].ptr, bufs[1].ptr, grids[0].size, bufs[2].ptr, grids[1].size, bufs[3].ptr,
            grids[2].size, bufs[4].ptr, C.c_int64(n), out.ptr))
        renderer.synchronize()
        return out.to_host(np.empty((n, 11)))
    finally:
        for b in bufs + [out]:
            b.free()


@pytest.mark.parametrize("n", [1 << 16, 300_001])
def test_bucketed_sampler_equals_direct_gather(renderer, n):
    """From 2^16 points (and at least 32 per block of brick cells) the device sampler partitions the points by
    block before gathering; each point's arithmetic is unchanged, so the result equals the unsorted gather (calls below
    the threshold) bit for bit, NaN rows outside the r/theta range included, and the oracle within the
    reference's 1e-12 gate."""
    print(f"bucketed sampler n={n}", flush=True)
    rng = np.random.default_rng(n)
    br = scenes.make_brick(20, 14, 18, spin=0.9, seed=3)
    coeffs = br.coeffs.astype(np.float64)
    grids = [g.astype(np.float64) for g in (br.r, br.theta, br.phi)]
    pts = np.empty((n, 3))
    span_r, span_t = grids[0][-1] - grids[0][0], grids[1][-1] - grids[1][0]
    pts[:, 0] = rng.uniform(grids[0][0] - 0.05 * span_r, grids[0][-1] + 0.05 * span_r, n)
    pts[:, 1] = rng.uniform(grids[1][0] - 0.05 * span_t, grids[1][-1] + 0.05 * span_t, n)
    pts[:, 2] = rng.uniform(-7.0, 13.0, n)
    pts[:64, 0] = grids[0][rng.integers(0, 20, 64)]   # exact grid values, the last one included
    got = _device_sample(renderer, coeffs, grids, pts)
    direct = np.concatenate([_device_sample(renderer, coeffs, grids, pts[a:a + 60000]) for a in range(0, n, 60000)])
    np.testing.assert_array_equal(got, direct)
    outside = np.isnan(got[:, 0])
    assert 0 < outside.mean() < 0.2
    pick = rng.choice(n, 5000, replace=False)
    ref, _ = micro.sample_vec(coeffs, *grids, pts[pick, 0], pts[pick, 1], pts[pick, 2])
    np.testing.assert_allclose(got[pick], ref, rtol=1e-12, atol=1e-12, equal_nan=True)


def test_partitioned_sampler_on_the_scale_brick(renderer):
    """The 128x64x128 brick (8192 blocks of 4x4x8 cells) at 2^20 points: grid values on every block face,
    the last r and theta nodes, phi on the seam and beyond 2 pi, 5 % outside; equal to the unsorted gather bit
    for bit and to the oracle within the reference's gate."""
    n = 1 << 20
    print(f"partitioned sampler, scale brick, n={n}", flush=True)
    rng = np.random.default_rng(77)
    br = scenes.scale_brick()
    coeffs = br.coeffs.astype(np.float64)
    grids = [g.astype(np.float64) for g in (br.r, br.theta, br.phi)]
    pts = np.empty((n, 3))
    span_r, span_t = grids[0][-1] - grids[0][0], grids[1][-1] - grids[1][0]
    pts[:, 0] = rng.uniform(grids[0][0] - 0.025 * span_r, grids[0][-1] + 0.025 * span_r, n)
    pts[:, 1] = rng.uniform(grids[1][0] - 0.025 * span_t, grids[1][-1] + 0.025 * span_t, n)
    pts[:, 2] = rng.uniform(-2.0 * np.pi, 4.0 * np.pi, n)
    k = 4096
    pts[:k, 0] = grids[0][rng.integers(0, grids[0].size, k)]
    pts[k:2 * k, 1] = grids[1][rng.integers(0, grids[1].size, k)]
    pts[2 * k:3 * k, 2] = grids[2][rng.integers(0, grids[2].size, k)] + 2.0 * np.pi * rng.integers(-1, 2, k)
    pts[3 * k:3 * k + 64, 2] = grids[2][-1]
    got = _device_sample(renderer, coeffs, grids, pts)
    direct = np.concatenate([_device_sample(renderer, coeffs, grids, pts[a:a + 60000]) for a in range(0, n, 60000)])
    np.testing.assert_array_equal(got, direct)
    assert 0.02 < np.isnan(got[:, 0]).mean() < 0.1
    pick = np.concatenate([np.arange(3 * k + 64), rng.choice(n, 4000, replace=False)])
    ref, _ = micro.sample_vec(coeffs, *grids, pts[pick, 0], pts[pick, 1], pts[pick, 2])
    np.testing.assert_allclose(got[pick], ref, rtol=1e-12, atol=1e-12, equal_nan=True)
