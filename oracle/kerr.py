"""FP64 oracle for Kerr geometry, the Hamiltonian RHS and the ZAMO camera (test infrastructure only).

Restates python/blackhole_sim/kerr.py in the reference's own matrix form:
g^{mu nu} (kerr.py:86-102), the analytic quotient-rule derivatives
d_r g^{mu nu}, d_theta g^{mu nu} (kerr.py:236-282) and
dx/dlambda = g p, dp/dlambda = -1/2 p (dg) p (kerr.py:285-303).  The HIP
kernels use the separated Carter form of the same Hamiltonian instead, so the
two are independent derivations checked against each other.
"""

from __future__ import annotations

import math

import numpy as np

EPS_THETA = 1.0e-7   # kerr.py:27


def validate_spin(a: float) -> None:
    """kerr.py:30-32."""
    if not np.isfinite(a) or abs(a) >= 1.0:
        raise ValueError("Kerr spin must be finite and satisfy |a| < 1 for this integrator")


def safe_theta(theta):
    """kerr.py:35-36 (works on scalars and arrays)."""
    return np.clip(theta, EPS_THETA, math.pi - EPS_THETA)


def horizon_radius(a: float) -> float:
    """kerr.py:39-42."""
    validate_spin(a)
    return 1.0 + math.sqrt(1.0 - a * a)


def isco_radius(a: float, prograde: bool = True) -> float:
    """kerr.py:387-394 (Bardeen-Press-Teukolsky)."""
    validate_spin(a)
    aa = abs(a)
    z1 = 1.0 + (1.0 - aa * aa) ** (1.0 / 3.0) * ((1.0 + aa) ** (1.0 / 3.0) + (1.0 - aa) ** (1.0 / 3.0))
    z2 = math.sqrt(3.0 * aa * aa + z1 * z1)
    sign = -1.0 if prograde else 1.0
    return 3.0 + z2 + sign * math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))


def _sigma(r, th, a):
    return r * r + a * a * np.cos(th) ** 2


def _delta(r, a):
    return r * r - 2.0 * r + a * a


def _big_a(r, th, a):
    return (r * r + a * a) ** 2 - a * a * _delta(r, a) * np.sin(th) ** 2


def metric_covariant(r: float, theta: float, a: float) -> np.ndarray:
    """g_{mu nu}, kerr.py:67-83."""
    validate_spin(a)
    th = float(safe_theta(theta))
    sig = _sigma(r, th, a)
    dlt = _delta(r, a)
    s2 = math.sin(th) ** 2
    A = _big_a(r, th, a)
    g = np.zeros((4, 4))
    g[0, 0] = -(1.0 - 2.0 * r / sig)
    g[0, 3] = g[3, 0] = -2.0 * a * r * s2 / sig
    g[1, 1] = sig / dlt
    g[2, 2] = sig
    g[3, 3] = A * s2 / sig
    return g


def metric_contravariant(r: float, theta: float, a: float) -> np.ndarray:
    """g^{mu nu}, kerr.py:86-102 (sin^2 floored at EPS_THETA^2)."""
    validate_spin(a)
    th = float(safe_theta(theta))
    sig = _sigma(r, th, a)
    dlt = _delta(r, a)
    s = math.sin(th)
    s2 = max(s * s, EPS_THETA * EPS_THETA)
    A = _big_a(r, th, a)
    g = np.zeros((4, 4))
    g[0, 0] = -A / (sig * dlt)
    g[0, 3] = g[3, 0] = -2.0 * a * r / (sig * dlt)
    g[1, 1] = dlt / sig
    g[2, 2] = 1.0 / sig
    g[3, 3] = (dlt - a * a * s2) / (sig * dlt * s2)
    return g


def hamiltonian_null(y, a: float, p_t: float, p_phi: float) -> float:
    """kerr.py:113-121; y = [t, r, theta, phi, p_r, p_theta]."""
    _, r, th, _, p_r, p_th = y
    p = np.array([p_t, p_r, p_th, p_phi])
    return float(0.5 * p @ metric_contravariant(float(r), float(th), a) @ p)


def zamo_tetrad(r: float, theta: float, a: float) -> np.ndarray:
    """kerr.py:124-152: rows e_(t), e_(r), e_(theta), e_(phi) in BL components."""
    validate_spin(a)
    th = float(safe_theta(theta))
    sig = _sigma(r, th, a)
    dlt = _delta(r, a)
    A = _big_a(r, th, a)
    if dlt <= 0.0:
        raise ValueError("ZAMO tetrad must be outside the outer horizon")
    s = math.sin(th)
    gcov = metric_covariant(r, th, a)
    lapse = math.sqrt(sig * dlt / A)
    omega = 2.0 * a * r / A
    e = np.zeros((4, 4))
    e[0] = [1.0 / lapse, 0.0, 0.0, omega / lapse]
    e[1] = [0.0, math.sqrt(dlt / sig), 0.0, 0.0]
    e[2] = [0.0, 0.0, 1.0 / math.sqrt(sig), 0.0]
    e[3] = [0.0, 0.0, 0.0, 1.0 / math.sqrt(gcov[3, 3])]
    if abs(s) < EPS_THETA:
        raise ValueError("ZAMO tetrad is ill-conditioned at the coordinate pole")
    return e


def camera_constants(camera, a: float, aspect: float):
    """Per-render constants of camera_ray_initial_state (kerr.py:208-232): tan_x, tan_y, tetrad, g_cov, theta."""
    th = float(safe_theta(camera.theta))
    tan_y = math.tan(0.5 * math.radians(camera.fov_y_degrees))
    tan_x = aspect * tan_y
    return tan_x, tan_y, zamo_tetrad(camera.r, th, a), metric_covariant(camera.r, th, a), th


def camera_ray_initial_state(camera, a: float, ndc_x: float, ndc_y: float, aspect: float):
    """kerr.py:188-233: returns (y0, p_t, p_phi) with y0 = [t, r, theta, phi, p_r, p_theta]."""
    tan_x, tan_y, tetrad, gcov, th = camera_constants(camera, a, aspect)
    n_r = -1.0
    n_theta = -ndc_y * tan_y
    n_phi = ndc_x * tan_x
    if camera.roll_degrees:
        roll = math.radians(camera.roll_degrees)
        ct, st = math.cos(roll), math.sin(roll)
        n_theta, n_phi = ct * n_theta - st * n_phi, st * n_theta + ct * n_phi
    n = np.array([n_r, n_theta, n_phi])
    n /= np.linalg.norm(n)
    p_con = tetrad[0] + n[0] * tetrad[1] + n[1] * tetrad[2] + n[2] * tetrad[3]
    p_cov = gcov @ p_con
    y0 = np.array([0.0, camera.r, th, camera.phi, p_cov[1], p_cov[2]])
    return y0, float(p_cov[0]), float(p_cov[3])


def camera_rays_vec(camera, a: float, ndc_x: np.ndarray, ndc_y: np.ndarray, aspect: float):
    """Vectorised camera_ray_initial_state over pixels: returns (r, th, ph, p_r, p_th, p_t, p_phi) arrays."""
    tan_x, tan_y, e, g, th = camera_constants(camera, a, aspect)
    n_r = -np.ones_like(ndc_x)
    n_theta = -ndc_y * tan_y
    n_phi = ndc_x * tan_x
    if camera.roll_degrees:
        roll = math.radians(camera.roll_degrees)
        ct, st = math.cos(roll), math.sin(roll)
        n_theta, n_phi = ct * n_theta - st * n_phi, st * n_theta + ct * n_phi
    norm = np.sqrt(n_r * n_r + n_theta * n_theta + n_phi * n_phi)
    n_r, n_theta, n_phi = n_r / norm, n_theta / norm, n_phi / norm
    pc_t = e[0, 0] + n_r * e[1, 0] + n_theta * e[2, 0] + n_phi * e[3, 0]
    pc_r = e[0, 1] + n_r * e[1, 1] + n_theta * e[2, 1] + n_phi * e[3, 1]
    pc_th = e[0, 2] + n_r * e[1, 2] + n_theta * e[2, 2] + n_phi * e[3, 2]
    pc_ph = e[0, 3] + n_r * e[1, 3] + n_theta * e[2, 3] + n_phi * e[3, 3]
    p_t = g[0, 0] * pc_t + g[0, 3] * pc_ph
    p_r = g[1, 1] * pc_r
    p_th = g[2, 2] * pc_th
    p_ph = g[3, 0] * pc_t + g[3, 3] * pc_ph
    ones = np.ones_like(ndc_x)
    return camera.r * ones, th * ones, camera.phi * ones, p_r, p_th, p_t, p_ph


def metric_derivative(r: float, theta: float, a: float, wrt: str) -> np.ndarray:
    """kerr.py:236-282: analytic d g^{mu nu} / d r or d theta by the quotient rule."""
    th = float(safe_theta(theta))
    s, c = math.sin(th), math.cos(th)
    s2 = max(s * s, EPS_THETA * EPS_THETA)
    ds2 = 2.0 * s * c
    sig = _sigma(r, th, a)
    dlt = _delta(r, a)
    A = _big_a(r, th, a)
    if wrt == "r":
        sig_p, dlt_p, A_p, s2_p = 2.0 * r, 2.0 * r - 2.0, 4.0 * r * (r * r + a * a) - a * a * (2.0 * r - 2.0) * s2, 0.0
    else:
        sig_p, dlt_p, A_p, s2_p = -a * a * ds2, 0.0, -a * a * dlt * ds2, ds2
    out = np.zeros((4, 4))
    den = sig * dlt
    den_p = sig_p * dlt + sig * dlt_p
    out[0, 0] = -((A_p * den - A * den_p) / (den * den))
    num = -2.0 * a * r
    num_p = -2.0 * a if wrt == "r" else 0.0
    out[0, 3] = out[3, 0] = (num_p * den - num * den_p) / (den * den)
    out[1, 1] = (dlt_p * sig - dlt * sig_p) / (sig * sig)
    out[2, 2] = -sig_p / (sig * sig)
    num = dlt - a * a * s2
    num_p = dlt_p - a * a * s2_p
    den = sig * dlt * s2
    den_p = sig_p * dlt * s2 + sig * dlt_p * s2 + sig * dlt * s2_p
    out[3, 3] = (num_p * den - num * den_p) / (den * den)
    return out


def geodesic_rhs(y, a: float, p_t: float, p_phi: float) -> np.ndarray:
    """kerr.py:285-303: d/dlambda of [t, r, theta, phi, p_r, p_theta]."""
    _, r, th, _, p_r, p_th = y
    th = float(safe_theta(float(th)))
    p = np.array([p_t, p_r, p_th, p_phi])
    xdot = metric_contravariant(float(r), th, a) @ p
    dp_r = -0.5 * float(p @ metric_derivative(float(r), th, a, "r") @ p)
    dp_th = -0.5 * float(p @ metric_derivative(float(r), th, a, "theta") @ p)
    return np.array([xdot[0], xdot[1], xdot[2], xdot[3], dp_r, dp_th])


def rk2_step(y, h: float, a: float, p_t: float, p_phi: float) -> np.ndarray:
    """RK2 midpoint (kerr_stokes_kernel.hip:13 scheme; theta clamp of kerr.py:312)."""
    k1 = geodesic_rhs(y, a, p_t, p_phi)
    k2 = geodesic_rhs(y + 0.5 * h * k1, a, p_t, p_phi)
    out = y + h * k2
    out[2] = float(safe_theta(float(out[2])))
    return out


def rk4_step(y, h: float, a: float, p_t: float, p_phi: float) -> np.ndarray:
    """kerr.py:306-313: classical RK4, out = y + (h / 6) (k1 + 2 k2 + 2 k3 + k4), theta clamped."""
    k1 = geodesic_rhs(y, a, p_t, p_phi)
    k2 = geodesic_rhs(y + 0.5 * h * k1, a, p_t, p_phi)
    k3 = geodesic_rhs(y + 0.5 * h * k2, a, p_t, p_phi)
    k4 = geodesic_rhs(y + h * k3, a, p_t, p_phi)
    out = y + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    out[2] = float(safe_theta(float(out[2])))
    return out


# --- vectorised RHS over rays (same matrix form, component-wise) -----------

def geodesic_rhs_vec(r, th, p_r, p_th, p_t, p_ph, a: float):
    """Vectorised kerr.py:285-303 over arrays; returns (dr, dth, dph, dp_r, dp_th).  dt is not needed."""
    th = safe_theta(th)
    s, c = np.sin(th), np.cos(th)
    s2 = np.maximum(s * s, EPS_THETA * EPS_THETA)
    s2raw = s * s
    ds2 = 2.0 * s * c
    aa = a * a
    sig = r * r + aa * c * c
    dlt = r * r - 2.0 * r + aa
    A = (r * r + aa) ** 2 - aa * dlt * s2raw
    # g^{mu nu}
    g03 = -2.0 * a * r / (sig * dlt)
    g11 = dlt / sig
    g22 = 1.0 / sig
    g33 = (dlt - aa * s2) / (sig * dlt * s2)
    dr = g11 * p_r
    dth = g22 * p_th
    dph = g03 * p_t + g33 * p_ph

    def quad(sig_p, dlt_p, A_p, s2_p, num_p_tphi):
        den = sig * dlt
        den_p = sig_p * dlt + sig * dlt_p
        d00 = -((A_p * den - A * den_p) / (den * den))
        d03 = (num_p_tphi * den - (-2.0 * a * r) * den_p) / (den * den)
        d11 = (dlt_p * sig - dlt * sig_p) / (sig * sig)
        d22 = -sig_p / (sig * sig)
        num = dlt - aa * s2
        num_p = dlt_p - aa * s2_p
        den3 = sig * dlt * s2
        den3_p = sig_p * dlt * s2 + sig * dlt_p * s2 + sig * dlt * s2_p
        d33 = (num_p * den3 - num * den3_p) / (den3 * den3)
        return (p_t * d00 + p_ph * d03) * p_t + p_r * d11 * p_r + p_th * d22 * p_th + (p_t * d03 + p_ph * d33) * p_ph

    zero = np.zeros_like(r)
    q_r = quad(2.0 * r, 2.0 * r - 2.0, 4.0 * r * (r * r + aa) - aa * (2.0 * r - 2.0) * s2, zero, -2.0 * a + zero)
    q_th = quad(-aa * ds2, zero, -aa * dlt * ds2, ds2, zero)
    return dr, dth, dph, -0.5 * q_r, -0.5 * q_th


def hamiltonian_vec(r, th, p_r, p_th, p_t, p_ph, a: float):
    """Vectorised kerr.py:113-121."""
    th = safe_theta(th)
    s, c = np.sin(th), np.cos(th)
    s2 = np.maximum(s * s, EPS_THETA * EPS_THETA)
    aa = a * a
    sig = r * r + aa * c * c
    dlt = r * r - 2.0 * r + aa
    A = (r * r + aa) ** 2 - aa * dlt * s * s
    g00 = -A / (sig * dlt)
    g03 = -2.0 * a * r / (sig * dlt)
    g33 = (dlt - aa * s2) / (sig * dlt * s2)
    return 0.5 * (g00 * p_t * p_t + 2.0 * g03 * p_t * p_ph + dlt / sig * p_r * p_r + p_th * p_th / sig + g33 * p_ph * p_ph)


def carter_constant_vec(th, p_th, p_t, p_ph, a: float):
    """Carter constant Q = p_theta^2 + cos^2 theta (-a^2 p_t^2 + p_phi^2 / sin^2 theta) (textbook)."""
    c, s = np.cos(th), np.sin(th)
    return p_th * p_th + c * c * (-a * a * p_t * p_t + p_ph * p_ph / (s * s))


def critical_curve(a: float, inclination: float, n: int = 2000):
    """Bardeen's closed-form Kerr shadow boundary (alpha, beta) for a distant observer at inclination i.

    Spherical photon orbits at radius r_p in [r_ph+, r_ph-] have
    xi = (r^2 (3 - r) - a^2 (r + 1)) / (a (r - 1)),
    eta = r^3 (4 a^2 - r (r - 3)^2) / (a^2 (r - 1)^2);
    alpha = -xi / sin i, beta = +-sqrt(eta + a^2 cos^2 i - xi^2 cot^2 i).
    Returns the closed curve as arrays (alpha, beta) over the visible part.
    """
    r1 = 2.0 * (1.0 + math.cos(2.0 / 3.0 * math.acos(-abs(a))))
    r2 = 2.0 * (1.0 + math.cos(2.0 / 3.0 * math.acos(abs(a))))
    rp = np.linspace(r2, r1, n)
    aa = a if abs(a) > 1e-9 else 1e-9
    xi = (rp * rp * (3.0 - rp) - aa * aa * (rp + 1.0)) / (aa * (rp - 1.0))
    eta = rp ** 3 * (4.0 * aa * aa - rp * (rp - 3.0) ** 2) / (aa * aa * (rp - 1.0) ** 2)
    si, ci = math.sin(inclination), math.cos(inclination)
    b2 = eta + aa * aa * ci * ci - xi * xi * (ci / si) ** 2
    ok = b2 >= 0.0
    alpha = -xi[ok] / si
    beta = np.sqrt(b2[ok])
    return np.concatenate([alpha, alpha[::-1]]), np.concatenate([beta, -beta[::-1]])
