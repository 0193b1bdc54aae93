"""FP64 oracle for the coefficient-brick precompute (test infrastructure only).

Restates python/blackhole_sim/synchrotron.py (the hybrid thermal + power-law synchrotron model,
:122-325) and python/blackhole_sim/coefficient_bricks.py:93-152 (cell frames and
``precompute_coefficient_bricks``), vectorised over cells: every cell runs the reference's scalar
recipe, with NumPy doing the same operations on arrays.  scipy.special.kv supplies K_2 and K_{2/3}
as in the reference.  The analytic torus fixture the reference's tests use is restated by
``bhgrrt.grmhd``; this module takes the snapshot fields as arrays.

Pinned by tests/golden/synchrotron_vectors.npz (written by the reference itself), and by closed forms:
Kirchhoff's law for the thermal part, the EVPA rotation invariants, linearity in the electron
density and the Faraday-conversion formula.
"""

from __future__ import annotations

import math

import numpy as np
from scipy import special

# calibration.py:18-26 (cgs)
C_CGS = 2.99792458e10
M_PROTON_CGS = 1.67262192369e-24
M_ELECTRON_CGS = 9.1093837015e-28
K_BOLTZMANN_CGS = 1.380649e-16
E_CHARGE_ESU = 4.80320471257e-10
H_CGS = 6.62607015e-27   # synchrotron.py:200

NAMES = ("j_i", "j_q", "j_u", "j_v", "alpha_i", "alpha_q", "alpha_u", "alpha_v", "rho_v", "rho_q", "rho_u")


def thermal_params(gamma_grid_size=192, pitch_nodes=16, faraday_rotation_scale=1.0, faraday_conversion_scale=1.0):
    """ThermalSynchrotronCoefficients defaults (synchrotron.py:231-238)."""
    return dict(gamma_grid_size=gamma_grid_size, pitch_nodes=pitch_nodes,
                faraday_rotation_scale=faraday_rotation_scale, faraday_conversion_scale=faraday_conversion_scale)


def nonthermal_params(p=3.5, gamma_min=10.0, gamma_max=1.0e5, nonthermal_fraction=0.02, gamma_grid_size=224,
                      pitch_nodes=16, absorption_scale=1.0, faraday_scale=0.15):
    """NonthermalPowerLawSynchrotronCoefficients defaults (synchrotron.py:271-282)."""
    return dict(p=p, gamma_min=gamma_min, gamma_max=gamma_max, nonthermal_fraction=nonthermal_fraction,
                gamma_grid_size=gamma_grid_size, pitch_nodes=pitch_nodes, absorption_scale=absorption_scale,
                faraday_scale=faraday_scale)


# --- kernel table (synchrotron.py:122-171) -----------------------------------------

def f_approx(x):
    """synchrotron.py:122-138: 1.25 x^(1/3) e^-x (648 + x^2)^(1/12) on x clipped to [1e-12, 1e3]."""
    xx = np.clip(np.asarray(x, dtype=float), 1.0e-12, 1.0e3)
    out = 1.25 * np.power(xx, 1.0 / 3.0) * np.exp(-xx) * np.power(648.0 + xx * xx, 1.0 / 12.0)
    return np.nan_to_num(out, nan=0.0, posinf=0.0, neginf=0.0)


def g_kernel(x):
    """synchrotron.py:141-155: x K_{2/3}(x) (its non-finite fallbacks never fire on [1e-12, 1e3])."""
    xx = np.clip(np.asarray(x, dtype=float), 1.0e-12, 1.0e3)
    out = xx * special.kv(2.0 / 3.0, xx)
    assert np.all(np.isfinite(out))
    return np.nan_to_num(out, nan=0.0, posinf=0.0, neginf=0.0)


def kernel_table(n=256):
    """synchrotron.py:158-163."""
    xs = np.geomspace(1.0e-8, 1.0e3, n)
    return xs, f_approx(xs), g_kernel(xs)


def kernel_interp(x, table=None):
    """synchrotron.py:166-174: log-log interpolation of F and G."""
    xs, F, G = table or kernel_table()
    lx = np.log(np.clip(np.asarray(x, dtype=float), xs[0], xs[-1]))
    lxs = np.log(xs)
    outF = np.exp(np.interp(lx, lxs, np.log(np.maximum(F, 1.0e-300))))
    outG = np.exp(np.interp(lx, lxs, np.log(np.maximum(G, 1.0e-300))))
    return outF, outG


# --- distributions and integrals (synchrotron.py:177-228) ------------------------------

def planck_nu(nu_hz, theta_e):
    """synchrotron.py:177-185, per cell."""
    out = np.empty_like(theta_e)
    for i, th in enumerate(np.asarray(theta_e, dtype=float).ravel()):
        T = float(th) * M_ELECTRON_CGS * C_CGS * C_CGS / K_BOLTZMANN_CGS
        x = H_CGS * nu_hz / max(K_BOLTZMANN_CGS * T, 1.0e-300)
        if x < 1.0e-3:
            out.flat[i] = 2.0 * nu_hz * nu_hz * K_BOLTZMANN_CGS * T / (C_CGS * C_CGS)
        else:
            out.flat[i] = 2.0 * H_CGS * nu_hz**3 / (C_CGS * C_CGS) / max(math.expm1(min(x, 700.0)), 1.0e-300)
    return out


def thermal_distribution(gamma, theta_e):
    """synchrotron.py:192-199; gamma (cells, G), theta_e (cells,)."""
    beta = np.sqrt(np.maximum(1.0 - 1.0 / np.maximum(gamma, 1.0) ** 2, 0.0))
    th = np.maximum(theta_e, 1.0e-12)[:, None]
    norm = theta_e[:, None] * special.kv(2.0, 1.0 / th)
    f = gamma * gamma * beta * np.exp(-gamma / th) / np.maximum(norm, 1.0e-300)
    integ = np.trapezoid(f, gamma, axis=-1)[:, None]
    return f / np.maximum(integ, 1.0e-300)


def powerlaw_distribution(gamma, p, gamma_min, gamma_max):
    """synchrotron.py:202-207; gamma (G,)."""
    mask = (gamma >= gamma_min) & (gamma <= gamma_max)
    f = np.zeros_like(gamma)
    f[mask] = gamma[mask] ** (-p)
    integ = np.trapezoid(f, gamma)
    return f / max(integ, 1.0e-300)


def integrated_j(n_e, b, nu_hz, gamma, f, pitch_nodes, table=None):
    """synchrotron.py:210-228: (j_I, j_Q) for every cell; gamma and f are (cells, G) or (G,)."""
    mu, w = np.polynomial.legendre.leggauss(pitch_nodes)
    n = n_e.size
    gamma = np.broadcast_to(gamma, (n, gamma.shape[-1]))
    f = np.broadcast_to(f, (n, f.shape[-1]))
    tot_i, tot_q = np.zeros(n), np.zeros(n)
    nu_b = E_CHARGE_ESU * b / (2.0 * math.pi * M_ELECTRON_CGS * C_CGS)
    for m, ww in zip(mu, w):
        sin_a = math.sqrt(max(0.0, 1.0 - float(m) ** 2))
        bp = b * sin_a
        nu_c = 1.5 * gamma * gamma * nu_b[:, None] * sin_a
        x = nu_hz / np.maximum(nu_c, 1.0e-300)
        F, G = kernel_interp(x, table)
        pref = (math.sqrt(3.0) * E_CHARGE_ESU**3 * bp / (M_ELECTRON_CGS * C_CGS**2))[:, None]
        tot_i += 0.5 * ww * np.trapezoid(f * (pref * F), gamma, axis=-1)
        tot_q += 0.5 * ww * np.trapezoid(f * (pref * G), gamma, axis=-1)
    ok = (n_e > 0.0) & (b > 0.0)
    return np.where(ok, n_e * tot_i / (4.0 * math.pi), 0.0), np.where(ok, n_e * tot_q / (4.0 * math.pi), 0.0)


def _rotated(c, evpa):
    """PolarizedCoefficients.rotated (synchrotron.py:68-79); c is (cells, 11)."""
    co, si = np.cos(2.0 * evpa), np.sin(2.0 * evpa)
    out = c.copy()
    for q, u in ((1, 2), (5, 6), (9, 10)):
        out[:, q] = c[:, q] * co - c[:, u] * si
        out[:, u] = c[:, q] * si + c[:, u] * co
    return out


def thermal_coefficients(n_e, theta_e, b, evpa, nu_hz, prm=None, table=None):
    """ThermalSynchrotronCoefficients.coefficients (synchrotron.py:240-266), cos(pitch to LOS) = 0."""
    prm = prm or thermal_params()
    n = n_e.size
    out = np.zeros((n, 11))
    live = (n_e > 0.0) & (b > 0.0) & (nu_hz > 0.0)
    if not live.any():
        return out
    ne, th, bb, ev = n_e[live], theta_e[live], b[live], evpa[live]
    gmax = np.maximum(10.0, 1.0 + 45.0 * th)
    gamma = 1.0 + np.geomspace(np.full_like(gmax, 1.0e-5), gmax - 1.0, prm["gamma_grid_size"], axis=-1)
    f = thermal_distribution(gamma, th)
    j_i, j_q = integrated_j(ne, bb, nu_hz, gamma, f, prm["pitch_nodes"], table)
    bnu = planck_nu(nu_hz, th)
    a_i = j_i / np.maximum(bnu, 1.0e-300)
    a_q = j_q / np.maximum(bnu, 1.0e-300)
    rel = np.maximum(1.0 / (th * th + 1.0), 1.0e-8)
    b_par = bb * 0.0
    b_perp = np.abs(bb) * math.sqrt(max(0.0, 1.0 - 0.0**2))
    rho_v = prm["faraday_rotation_scale"] * (E_CHARGE_ESU**3 / (math.pi * M_ELECTRON_CGS**2 * C_CGS**2))
    rho_v = rho_v * (ne * b_par / max(nu_hz * nu_hz, 1.0e-300) * rel)
    rho_q = prm["faraday_conversion_scale"] * (E_CHARGE_ESU**4 / (4.0 * math.pi**2 * M_ELECTRON_CGS**3 * C_CGS**3))
    rho_q = rho_q * (ne * b_perp**2 / max(nu_hz**3, 1.0e-300) * np.maximum(th, 1.0e-8))
    c = np.zeros((ne.size, 11))
    c[:, 0], c[:, 1] = np.maximum(j_i, 0.0), np.maximum(j_q, 0.0)
    c[:, 4], c[:, 5] = np.maximum(a_i, 0.0), np.maximum(a_q, 0.0)
    c[:, 8], c[:, 9] = rho_v, rho_q
    out[live] = _rotated(c, ev)
    return out


def nonthermal_coefficients(n_e, theta_e, b, evpa, nu_hz, prm=None, table=None):
    """NonthermalPowerLawSynchrotronCoefficients.coefficients (synchrotron.py:284-312), cos(pitch to LOS) = 0."""
    prm = prm or nonthermal_params()
    n = n_e.size
    out = np.zeros((n, 11))
    live = (n_e > 0.0) & (b > 0.0) & (nu_hz > 0.0) & (prm["nonthermal_fraction"] > 0.0)
    if not live.any():
        return out
    ne, bb, ev = n_e[live], b[live], evpa[live]
    gamma = np.geomspace(prm["gamma_min"], prm["gamma_max"], prm["gamma_grid_size"])
    f = powerlaw_distribution(gamma, prm["p"], prm["gamma_min"], prm["gamma_max"])
    sne = ne * prm["nonthermal_fraction"]
    j_i, j_q = integrated_j(sne, bb, nu_hz, gamma, f, prm["pitch_nodes"], table)
    b_perp = np.abs(bb) * math.sqrt(max(0.0, 1.0 - 0.0**2))
    b_par = bb * 0.0
    cg = np.maximum(prm["gamma_min"], np.sqrt(np.maximum(nu_hz / np.maximum(4.2e6 * b_perp, 1.0e-300), 1.0)))
    energy = cg * M_ELECTRON_CGS * C_CGS * C_CGS
    a_i = prm["absorption_scale"] * j_i / np.maximum(energy * nu_hz * nu_hz / (C_CGS * C_CGS), 1.0e-300)
    pol = min((prm["p"] + 1.0) / (prm["p"] + 7.0 / 3.0), 0.85)
    j_q = np.minimum(j_q, pol * j_i)
    a_q = np.minimum(a_i * pol, a_i)
    fs = prm["faraday_scale"]
    rho_v = fs * (E_CHARGE_ESU**3 / (math.pi * M_ELECTRON_CGS**2 * C_CGS**2))
    rho_v = rho_v * (sne * b_par / max(nu_hz * nu_hz, 1.0e-300) / max(prm["gamma_min"] ** 2, 1.0))
    rho_q = fs * (E_CHARGE_ESU**4 / (4.0 * math.pi**2 * M_ELECTRON_CGS**3 * C_CGS**3))
    rho_q = rho_q * (sne * b_perp**2 / max(nu_hz**3, 1.0e-300) * math.log(max(prm["gamma_max"] / prm["gamma_min"], 1.0)))
    c = np.zeros((ne.size, 11))
    c[:, 0], c[:, 1] = np.maximum(j_i, 0.0), np.maximum(j_q, 0.0)
    c[:, 4], c[:, 5] = np.maximum(a_i, 0.0), np.maximum(a_q, 0.0)
    c[:, 8], c[:, 9] = rho_v, rho_q
    out[live] = _rotated(c, ev)
    return out


def hybrid_coefficients(n_e, theta_e, b, evpa, nu_hz, thermal=None, nonthermal=None, table=None):
    """HybridSynchrotronCoefficients.coefficients (synchrotron.py:315-325): (cells, 11) in NAMES order."""
    n_e, theta_e, b, evpa = (np.atleast_1d(np.asarray(v, dtype=float)) for v in (n_e, theta_e, b, evpa))
    table = table or kernel_table()
    return (thermal_coefficients(n_e, theta_e, b, evpa, nu_hz, thermal, table)
            + nonthermal_coefficients(n_e, theta_e, b, evpa, nu_hz, nonthermal, table))


# --- cell frames and the brick (coefficient_bricks.py:93-152) -----------------------------

def fast_frames(rho, theta_e, b_con, rho_cgs_per_code, b_gauss_per_code, mean_molecular_weight_per_electron=1.0):
    """coefficient_bricks.py:110-116 with PhysicalScaling (calibration.py:113-119) for every cell:
    (n_e, theta_e, B, evpa), each flat over cells."""
    bv = np.asarray(b_con, dtype=float).reshape(-1, 4)[:, 1:]
    b_code = np.sqrt(np.sum(bv * bv, axis=-1))
    b_gauss = b_code * b_gauss_per_code
    n_e = np.asarray(rho, dtype=float).ravel() * rho_cgs_per_code / max(mean_molecular_weight_per_electron * M_PROTON_CGS,
                                                                          1.0e-300)
    evpa = np.where(b_code > 0.0, np.arctan2(bv[:, 2], bv[:, 1]), 0.0)
    th = np.maximum(np.asarray(theta_e, dtype=float).ravel(), 1.0e-8)
    return n_e, th, np.maximum(b_gauss, 0.0), evpa


def precompute_coefficient_bricks(rho, theta_e, b_con, rho_cgs_per_code, b_gauss_per_code, nu_hz=230e9,
                                  dtype="float32", stride=1, mean_molecular_weight_per_electron=1.0,
                                  thermal=None, nonthermal=None):
    """coefficient_bricks.py:118-152 on snapshot field arrays: (nr', nt', np', 11) in ``dtype``."""
    if stride < 1:
        raise ValueError("stride must be >= 1")
    sl = (slice(None, None, stride),) * 3
    rho, theta_e, b_con = np.asarray(rho)[sl], np.asarray(theta_e)[sl], np.asarray(b_con)[sl]
    shape = rho.shape
    n_e, th, b, evpa = fast_frames(rho, theta_e, b_con, rho_cgs_per_code, b_gauss_per_code,
                                   mean_molecular_weight_per_electron)
    c64 = hybrid_coefficients(n_e, th, b, evpa, nu_hz, thermal, nonthermal)
    dt = np.dtype(dtype)
    with np.errstate(over="ignore", invalid="ignore"):
        c = c64.astype(dt)
    c = np.nan_to_num(c, nan=0.0, posinf=np.finfo(dt).max, neginf=0.0)
    return c.reshape(shape + (11,))
