"""Pin the Kerr oracle (oracle/kerr.py) with the reference's tests and with invariants of the geometry."""

import math

import numpy as np

from oracle import kerr
from bhgrrt.scenes import Camera


def test_horizon_and_isco_kat():
    # test_kerr.py:19-24
    assert np.isclose(kerr.horizon_radius(0.0), 2.0)
    assert np.isclose(kerr.horizon_radius(0.9), 1.0 + math.sqrt(1.0 - 0.81))
    assert np.isclose(kerr.isco_radius(0.0, prograde=True), 6.0)
    assert kerr.isco_radius(0.9, prograde=True) < 3.0
    assert kerr.isco_radius(0.9, prograde=False) > 8.0


def test_metric_inverse_identity():
    # test_kerr.py:27-31
    for r, th, a in ((8.0, 1.1, 0.72), (2.3, 0.4, 0.94), (40.0, 2.9, -0.5)):
        assert np.allclose(kerr.metric_covariant(r, th, a) @ kerr.metric_contravariant(r, th, a), np.eye(4), atol=1e-11)


def test_camera_ray_is_null_and_vectorised_camera_agrees():
    # test_kerr.py:34-37
    cam = Camera.from_degrees(r=40.0, inclination_degrees=62.0, fov_y_degrees=30.0)
    y0, pt, pph = kerr.camera_ray_initial_state(cam, 0.6, 0.1, -0.2, 16 / 9)
    assert abs(kerr.hamiltonian_null(y0, 0.6, pt, pph)) < 1e-11
    cam2 = Camera.from_degrees(r=1000.0, inclination_degrees=17.0, fov_y_degrees=2.2, roll_degrees=12.0)
    rng = np.random.default_rng(0)
    nx, ny = rng.uniform(-1, 1, 50), rng.uniform(-1, 1, 50)
    r, th, ph, pr, pth, pt_v, pph_v = kerr.camera_rays_vec(cam2, 0.94, nx, ny, 1.0)
    for i in range(50):
        y, a_pt, a_pph = kerr.camera_ray_initial_state(cam2, 0.94, nx[i], ny[i], 1.0)
        np.testing.assert_allclose([pr[i], pth[i], pt_v[i], pph_v[i]], [y[4], y[5], a_pt, a_pph], rtol=1e-14, atol=1e-15)
        assert abs(kerr.hamiltonian_null(y, 0.94, a_pt, a_pph)) < 1e-11


def test_metric_derivative_matches_finite_differences():
    rng = np.random.default_rng(2)
    for _ in range(20):
        r, th, a = rng.uniform(1.8, 30), rng.uniform(0.2, 2.9), rng.uniform(-0.95, 0.95)
        if r <= kerr.horizon_radius(a) * 1.05:
            continue
        for wrt, hstep in (("r", 1e-4 * r), ("theta", 1e-4)):
            if wrt == "r":
                fd = (kerr.metric_contravariant(r + hstep, th, a) - kerr.metric_contravariant(r - hstep, th, a)) / (2 * hstep)
            else:
                fd = (kerr.metric_contravariant(r, th + hstep, a) - kerr.metric_contravariant(r, th - hstep, a)) / (2 * hstep)
            an = kerr.metric_derivative(r, th, a, wrt)
            np.testing.assert_allclose(an, fd, rtol=2e-6, atol=1e-9 * np.max(np.abs(an)))


def _H(r, th, pr, pth, pt, pph, a):
    p = np.array([pt, pr, pth, pph])
    return 0.5 * p @ kerr.metric_contravariant(r, th, a) @ p


def test_rhs_is_hamiltons_equations():
    rng = np.random.default_rng(4)
    for _ in range(20):
        a = rng.uniform(-0.95, 0.95)
        r, th = rng.uniform(kerr.horizon_radius(a) * 1.2, 30), rng.uniform(0.2, 2.9)
        pr, pth, pt, pph = rng.normal(), rng.normal(), -1.0, rng.normal() * 3
        d = kerr.geodesic_rhs(np.array([0.0, r, th, 0.0, pr, pth]), a, pt, pph)
        e = 1e-6
        dH_dpr = (_H(r, th, pr + e, pth, pt, pph, a) - _H(r, th, pr - e, pth, pt, pph, a)) / (2 * e)
        dH_dpth = (_H(r, th, pr, pth + e, pt, pph, a) - _H(r, th, pr, pth - e, pt, pph, a)) / (2 * e)
        dH_dpph = (_H(r, th, pr, pth, pt, pph + e, a) - _H(r, th, pr, pth, pt, pph - e, a)) / (2 * e)
        er = 1e-6 * r
        dH_dr = (_H(r + er, th, pr, pth, pt, pph, a) - _H(r - er, th, pr, pth, pt, pph, a)) / (2 * er)
        dH_dth = (_H(r, th + e, pr, pth, pt, pph, a) - _H(r, th - e, pr, pth, pt, pph, a)) / (2 * e)
        np.testing.assert_allclose(d[1:], [dH_dpr, dH_dpth, dH_dpph, -dH_dr, -dH_dth], rtol=1e-6, atol=1e-8)
        v = kerr.geodesic_rhs_vec(np.array([r]), np.array([th]), np.array([pr]), np.array([pth]), pt, pph, a)
        np.testing.assert_allclose([x[0] for x in v], d[[1, 2, 3, 4, 5]], rtol=1e-13, atol=1e-15)


def _trace(y0, pt, pph, a, h, n):
    y = y0.copy()
    for _ in range(n):
        y = kerr.rk2_step(y, h, a, pt, pph)
    return y


def test_rk2_conserves_hamiltonian_and_carter_constant_at_second_order():
    a = 0.7
    cam = Camera.from_degrees(r=30.0, inclination_degrees=70.0, fov_y_degrees=30.0)
    y0, pt, pph = kerr.camera_ray_initial_state(cam, a, 0.3, 0.2, 1.0)
    Q0 = kerr.carter_constant_vec(y0[2], y0[5], pt, pph, a)
    lam = 20.0
    errs = []
    for n in (400, 800):
        y = _trace(y0, pt, pph, a, lam / n, n)
        assert abs(kerr.hamiltonian_null(y, a, pt, pph)) < 1e-4
        errs.append(abs(kerr.carter_constant_vec(y[2], y[5], pt, pph, a) - Q0))
    assert errs[1] < 1e-4
    assert 3.0 < errs[0] / errs[1] < 5.0


def _radial_potential_captures(E, L, Q, a, r_from):
    """Ingoing photon is captured iff R(r) = ((r^2+a^2)E - aL)^2 - Delta (Q + (L - aE)^2) > 0 on [r_+, r_from]."""
    rp = kerr.horizon_radius(a)
    r = np.geomspace(rp, r_from, 400001)
    P2 = ((r * r + a * a) * E - a * L) ** 2
    R = P2 - (r * r - 2 * r + a * a) * (Q + (L - a * E) ** 2)
    return bool(np.all(R > 0)), float(np.min(R / P2))


def test_capture_escape_matches_radial_potential():
    """RK2 classification of rays from r=1000 agrees with the analytic turning-point criterion."""
    a = 0.94
    cam = Camera.spanning(r=1000.0, inclination_degrees=17.0, half_width_m=12.0, width=64, height=64)
    rng = np.random.default_rng(11)
    nx, ny = rng.uniform(-1, 1, 400), rng.uniform(-1, 1, 400)
    r, th, ph, pr, pth, pt, pph = kerr.camera_rays_vec(cam, a, nx, ny, 1.0)
    checked = 0
    for i in range(400):
        Q = kerr.carter_constant_vec(th[i], pth[i], pt[i], pph[i], a)
        cap, margin = _radial_potential_captures(-pt[i], pph[i], Q, a, r[i])
        if abs(margin) < 2e-2:   # too close to the critical curve for a finite-step check
            continue
        y = np.array([0.0, r[i], th[i], ph[i], pr[i], pth[i]])
        status = "max"
        rh = kerr.horizon_radius(a)
        prev = y[1]
        for _ in range(20000):
            y = kerr.rk2_step(y, 0.02 * (y[1] - rh), a, pt[i], pph[i])
            if not kerr.EPS_THETA < y[2] < math.pi - kerr.EPS_THETA:
                status = "pole"   # the theta clamp (kerr.py:35-36) fired: not a geodesic any more
                break
            if y[1] <= rh * (1 + 2e-4):
                status = "captured"
                break
            if y[1] > 220.0 and y[1] > prev:
                status = "escaped"
                break
            prev = y[1]
        if status == "pole":
            continue
        assert status == ("captured" if cap else "escaped"), (i, margin)
        checked += 1
        if checked >= 40:
            break
    assert checked >= 30


def test_schwarzschild_critical_impact_parameter():
    # test_geodesics.py:6-29: b_c = sqrt(27); b = 4.8 captured, b = 6.2 escapes
    for b, expect in ((4.8, "captured"), (6.2, "escaped")):
        r0 = 80.0
        sin_alpha = b * math.sqrt(1 - 2 / r0) / r0
        cam = Camera(r=r0, theta=math.pi / 2, phi=0.0, fov_y_degrees=90.0)
        tan_x = math.tan(math.radians(45.0))
        ndc_x = (sin_alpha / math.sqrt(1 - sin_alpha ** 2)) / tan_x
        y, pt, pph = kerr.camera_ray_initial_state(cam, 0.0, ndc_x, 0.0, 1.0)
        assert abs(pph / -pt - b) < 1e-9
        status, prev = "max", y[1]
        for _ in range(20000):
            y = kerr.rk2_step(y, 0.01 * (y[1] - 2.0), 0.0, pt, pph)
            if y[1] <= 2.0 * (1 + 2e-4):
                status = "captured"
                break
            if y[1] > 80.0 and y[1] > prev:
                status = "escaped"
                break
            prev = y[1]
        assert status == expect


def test_critical_curve_schwarzschild_limit():
    al, be = kerr.critical_curve(1e-6, math.radians(60.0))
    np.testing.assert_allclose(np.hypot(al, be), math.sqrt(27.0), rtol=1e-3)


def test_camera_orientation_kat():
    """kerr.py:200-201: x positive is local +phi, y positive is local -theta on the image plane."""
    cam = Camera.from_degrees(r=60.0, inclination_degrees=70.0, fov_y_degrees=40.0)
    for a in (0.0, 0.9, -0.6):
        for ndc_y in (0.4, -0.7):
            y0, pt, pph = kerr.camera_ray_initial_state(cam, a, 0.0, ndc_y, 1.5)
            dth = kerr.geodesic_rhs(y0, a, pt, pph)[2]
            assert np.sign(y0[5]) == -np.sign(ndc_y) and np.sign(dth) == -np.sign(ndc_y)
        if a == 0.0:   # no frame dragging: d phi / d lambda has the sign of ndc_x
            for ndc_x in (0.3, -0.8):
                y0, pt, pph = kerr.camera_ray_initial_state(cam, a, ndc_x, 0.1, 1.5)
                assert np.sign(pph) == np.sign(ndc_x)
                assert np.sign(kerr.geodesic_rhs(y0, a, pt, pph)[3]) == np.sign(ndc_x)


def _bardeen(pt, pph, th, pth, a, inc):
    """Bardeen's image coordinates of a ray from its constants: alpha = -xi / sin i,
    beta = sqrt(eta + a^2 cos^2 i - xi^2 cot^2 i) (xi = L / E, eta = Q / E^2)."""
    E, L = -pt, pph
    Q = kerr.carter_constant_vec(th, pth, pt, pph, a)
    xi, eta = L / E, Q / (E * E)
    b2 = eta + a * a * math.cos(inc) ** 2 - xi * xi / math.tan(inc) ** 2
    return -xi / math.sin(inc), math.sqrt(max(b2, 0.0))


def test_pixel_offsets_are_bardeen_image_coordinates():
    """A camera at r = 1000 sees a ray at image-plane offsets (r ndc_x tan_x, r ndc_y tan_y) M; these equal
    Bardeen's (-alpha, |beta|) from the ray's constants of motion to O(M / r): the image's +x is local
    +phi (xi = L / E > 0, alpha = -xi / sin i < 0), and beta's sign is opposite to p_theta's (the
    image's +y is -theta)."""
    r0 = 1000.0
    for a, inc_deg in ((0.94, 17.0), (0.998, 60.0), (-0.5, 80.0)):
        cam = Camera.spanning(r=r0, inclination_degrees=inc_deg, half_width_m=20.0, width=64, height=48)
        tan_y = math.tan(math.radians(cam.fov_y_degrees) / 2)
        tan_x = tan_y * 64 / 48
        rng = np.random.default_rng(5)
        for nx, ny in zip(rng.uniform(-1, 1, 25), rng.uniform(-1, 1, 25)):
            y0, pt, pph = kerr.camera_ray_initial_state(cam, a, nx, ny, 64 / 48)
            al, be = _bardeen(pt, pph, y0[2], y0[5], a, cam.theta)
            a_pix, b_pix = r0 * nx * tan_x, r0 * ny * tan_y
            assert abs(-al - a_pix) < 2e-3 * max(1.0, abs(a_pix)), (a, inc_deg, al, a_pix)
            assert abs(be - abs(b_pix)) < 2e-3 * max(1.0, abs(b_pix)) + 2e-2, (a, inc_deg, be, b_pix)
            if abs(b_pix) > 0.5:
                assert np.sign(b_pix) == -np.sign(y0[5])


def test_kerr_critical_curve_separates_capture_from_escape():
    """At a = 0.998, rays launched 0.05 M inside the closed-form critical curve are captured and rays 0.05 M
    outside escape, by the radial potential of each ray's own constants of motion (textbook turning-point
    criterion): the curve's position, including the flat prograde side, is pinned, not only its a -> 0 circle."""
    a, inc, r0 = 0.998, math.radians(60.0), 1000.0
    al, be = kerr.critical_curve(a, inc, n=4000)
    cam = Camera.spanning(r=r0, inclination_degrees=60.0, half_width_m=20.0, width=64, height=64)
    tan = math.tan(math.radians(cam.fov_y_degrees) / 2)
    centre = np.array([np.mean(al), np.mean(be)])
    picks = np.linspace(0, al.size - 1, 24).astype(int)
    checked = 0
    for k in picks:
        p = np.array([al[k], be[k]])
        if abs(p[1]) < 0.3:   # the curve's two halves meet at beta = 0, where the normal is least defined
            continue
        k0, k1 = max(k - 3, 0), min(k + 3, al.size - 1)
        t = np.array([al[k1] - al[k0], be[k1] - be[k0]])
        nrm = np.array([t[1], -t[0]]) / np.hypot(*t)
        if np.dot(nrm, p - centre) < 0:
            nrm = -nrm   # outward
        for sgn, expect in ((-1.0, True), (1.0, False)):
            q = p + sgn * 0.05 * nrm
            # Bardeen (alpha, beta) -> pixel offsets (-alpha, beta)
            y0, pt, pph = kerr.camera_ray_initial_state(cam, a, -q[0] / (r0 * tan), q[1] / (r0 * tan), 1.0)
            Q = kerr.carter_constant_vec(y0[2], y0[5], pt, pph, a)
            cap, _ = _radial_potential_captures(-pt, pph, Q, a, r0)
            assert cap == expect, (k, p, sgn)
        checked += 1
    assert checked >= 16
