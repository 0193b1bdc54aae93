"""FP64 oracle for the four drop-in micro-kernels (test infrastructure only).

Restates python/blackhole_sim/native_kernels.py and the Rust core
native/core/src/lib.rs.  Scalar functions follow the reference line by line;
the ``*_vec`` functions are NumPy-vectorised equivalents used for large parity
cases and are themselves checked against the scalar ones.
"""

from __future__ import annotations

import math

import numpy as np

TWO_PI = 2.0 * math.pi
RTOL = 1.0e-12   # native_kernels.py:12-15
ATOL = 1.0e-12
POLICIES = ("nan", "initial", "zero")   # native_kernels.py:16


# --- validation (native_kernels.py:53-130) ---------------------------------

def coerce_coefficients(coeffs):
    """native_kernels.py:53-61."""
    arr = np.asarray(coeffs, dtype=np.float64)
    if arr.ndim < 1 or arr.shape[-1] != 11:
        raise ValueError("coeffs must have shape (..., 11)")
    if arr.size == 0:
        raise ValueError("coeffs must contain at least one cell")
    if not np.all(np.isfinite(arr)):
        raise ValueError("coeffs must contain only finite values")
    return np.ascontiguousarray(arr), tuple(int(v) for v in arr.shape[:-1])


def coerce_grid(name, grid):
    """native_kernels.py:64-72."""
    g = np.asarray(grid, dtype=np.float64)
    if g.ndim != 1 or g.size < 2:
        raise ValueError(f"{name} grid must contain at least 2 values")
    if not np.all(np.isfinite(g)):
        raise ValueError(f"{name} grid must contain only finite values")
    if np.any(np.diff(g) <= 0.0):
        raise ValueError(f"{name} grid must be strictly increasing")
    return np.ascontiguousarray(g)


def coerce_points(points):
    """native_kernels.py:75-85."""
    p = np.asarray(points, dtype=np.float64)
    if p.ndim == 1:
        if p.size % 3 != 0:
            raise ValueError("points must contain r/theta/phi triples")
        p = p.reshape((-1, 3))
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError("points must have shape (n, 3) or be flat r/theta/phi triples")
    if not np.all(np.isfinite(p)):
        raise ValueError("points must contain only finite values")
    return np.ascontiguousarray(p)


def coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points):
    """native_kernels.py:88-104."""
    c, prefix = coerce_coefficients(coeffs)
    if len(prefix) != 3:
        raise ValueError("coeffs must have shape (r, theta, phi, 11)")
    rg = coerce_grid("r", r_grid)
    tg = coerce_grid("theta", theta_grid)
    pg = coerce_grid("phi", phi_grid)
    expect = (int(rg.size), int(tg.size), int(pg.size))
    if prefix != expect:
        raise ValueError(f"coeffs grid shape must be {expect}; got {prefix}")
    return c, rg, tg, pg, coerce_points(points)


def coerce_initial(initial, prefix):
    """native_kernels.py:107-117."""
    if initial is None:
        return np.zeros(tuple(prefix) + (4,), dtype=np.float64)
    arr = np.asarray(initial, dtype=np.float64)
    if not np.all(np.isfinite(arr)):
        raise ValueError("initial must contain only finite values")
    if arr.shape == (4,):
        return np.ascontiguousarray(np.broadcast_to(arr, tuple(prefix) + (4,)))
    if arr.shape != tuple(prefix) + (4,):
        raise ValueError(f"initial must have shape (4,) or {tuple(prefix) + (4,)}")
    return np.ascontiguousarray(arr)


def validate_ds(ds):
    """native_kernels.py:120-124."""
    v = float(ds)
    if not np.isfinite(v) or v < 0.0:
        raise ValueError("ds_cm must be finite and non-negative")
    return v


def coerce_policy(policy):
    """native_kernels.py:127-130."""
    if policy not in POLICIES:
        raise ValueError("invalid_policy must be 'nan', 'initial', or 'zero'")
    return policy


# --- brackets (native_kernels.py:133-161 == lib.rs:61-106) -----------------

def wrap_phi(phi: float, base: float = 0.0) -> float:
    """native_kernels.py:133-134: Python float % is a floor modulo."""
    return float(((phi - base) % TWO_PI) + base)


def search_right(grid: np.ndarray, x: float) -> int:
    """Number of grid values <= x (np.searchsorted side='right'; lib.rs:61-73)."""
    lo, hi = 0, int(grid.size)
    while lo < hi:
        mid = lo + (hi - lo) // 2
        if x < grid[mid]:
            hi = mid
        else:
            lo = mid + 1
    return lo


def bracket_linear(grid: np.ndarray, x: float):
    """native_kernels.py:137-146."""
    if x < grid[0] or x > grid[-1]:
        return None
    if x == grid[-1]:
        return int(grid.size - 2), int(grid.size - 1), 1.0
    i0 = search_right(grid, x) - 1
    i0 = max(0, min(i0, int(grid.size - 2)))
    i1 = i0 + 1
    w = (x - grid[i0]) / max(grid[i1] - grid[i0], 1.0e-300)
    return i0, i1, float(w)


def bracket_periodic_phi(grid: np.ndarray, phi: float):
    """native_kernels.py:149-161."""
    p = wrap_phi(phi, float(grid[0]))
    n = int(grid.size)
    i0 = search_right(grid, p) - 1
    if i0 < 0:
        i0 = n - 1
    if i0 >= n:
        i0 = n - 1
    i1 = (i0 + 1) % n
    hi = grid[i1] if i1 > i0 else grid[0] + TWO_PI
    pp = p if p >= grid[i0] else p + TWO_PI
    w = (pp - grid[i0]) / max(hi - grid[i0], 1.0e-300)
    return i0, i1, float(w)


def sample_one(coeffs, rg, tg, pg, point) -> np.ndarray:
    """native_kernels.py:164-186: phi, then theta, then r lerps of (1-w)a + wb."""
    rb = bracket_linear(rg, float(point[0]))
    tb = bracket_linear(tg, float(point[1]))
    if rb is None or tb is None:
        return np.full((11,), np.nan)
    r0, r1, wr = rb
    t0, t1, wt = tb
    p0, p1, wp = bracket_periodic_phi(pg, float(point[2]))
    c00 = (1.0 - wp) * coeffs[r0, t0, p0] + wp * coeffs[r0, t0, p1]
    c01 = (1.0 - wp) * coeffs[r0, t1, p0] + wp * coeffs[r0, t1, p1]
    c10 = (1.0 - wp) * coeffs[r1, t0, p0] + wp * coeffs[r1, t0, p1]
    c11 = (1.0 - wp) * coeffs[r1, t1, p0] + wp * coeffs[r1, t1, p1]
    c0 = (1.0 - wt) * c00 + wt * c01
    c1 = (1.0 - wt) * c10 + wt * c11
    return np.ascontiguousarray((1.0 - wr) * c0 + wr * c1)


def sample_brick_valid_mask(r_grid, theta_grid, points) -> np.ndarray:
    """native_kernels.py:189-210: r/theta inside the closed grid ranges; phi ignored."""
    rg = coerce_grid("r", r_grid)
    tg = coerce_grid("theta", theta_grid)
    p = coerce_points(points)
    return np.ascontiguousarray((p[:, 0] >= rg[0]) & (p[:, 0] <= rg[-1]) & (p[:, 1] >= tg[0]) & (p[:, 1] <= tg[-1]))


def stokes_rhs(stokes: np.ndarray, c: np.ndarray) -> np.ndarray:
    """native_kernels.py:213-222 == lib.rs:37-59: j - K S (K of synchrotron.py:82-94)."""
    out = np.empty_like(stokes, dtype=np.float64)
    out[:, 0] = c[:, 0] - (c[:, 4] * stokes[:, 0] + c[:, 5] * stokes[:, 1] + c[:, 6] * stokes[:, 2] + c[:, 7] * stokes[:, 3])
    out[:, 1] = c[:, 1] - (c[:, 5] * stokes[:, 0] + c[:, 4] * stokes[:, 1] + c[:, 8] * stokes[:, 2] - c[:, 10] * stokes[:, 3])
    out[:, 2] = c[:, 2] - (c[:, 6] * stokes[:, 0] - c[:, 8] * stokes[:, 1] + c[:, 4] * stokes[:, 2] + c[:, 9] * stokes[:, 3])
    out[:, 3] = c[:, 3] - (c[:, 7] * stokes[:, 0] + c[:, 10] * stokes[:, 1] - c[:, 9] * stokes[:, 2] + c[:, 4] * stokes[:, 3])
    return out


def propagation_matrix(c: np.ndarray) -> np.ndarray:
    """synchrotron.py:82-94: K of dS/ds = j - K S for (n, 11) coefficient rows, shape (n, 4, 4)."""
    c = np.asarray(c, dtype=np.float64)
    ai, aq, au, av, rv, rq, ru = (c[:, k] for k in range(4, 11))
    return np.stack([np.stack([ai, aq, au, av], -1), np.stack([aq, ai, rv, -ru], -1),
                     np.stack([au, -rv, ai, rq], -1), np.stack([av, ru, -rq, ai], -1)], -2)


def stokes_exact_step(stokes: np.ndarray, c: np.ndarray, ds: np.ndarray) -> np.ndarray:
    """polarized_transfer.py:51-70 on (n, 4) states with per-row ds: expm of the augmented matrix
    [[-K, j], [0, 0]] ds (scipy.linalg.expm, the reference's routine) applied to [S; 1]; rows with
    ds <= 0 keep S (a NaN ds goes through, as in the reference's ``if ds_cm <= 0.0``)."""
    from scipy.linalg import expm

    S = np.asarray(stokes, dtype=np.float64)
    ds = np.asarray(ds, dtype=np.float64)
    n = S.shape[0]
    A = np.zeros((n, 5, 5))
    A[:, :4, :4] = -propagation_matrix(c)
    A[:, :4, 4] = np.asarray(c, dtype=np.float64)[:, :4]
    with np.errstate(invalid="ignore", over="ignore"):
        M = expm(A * ds[:, None, None])
    y = np.concatenate([S, np.ones((n, 1))], axis=1)
    out = np.einsum("nij,nj->ni", M, y)[:, :4]
    return np.where((ds <= 0.0)[:, None], S, out)


def sample_brick_trilinear(coeffs, r_grid, theta_grid, phi_grid, points) -> np.ndarray:
    """native_kernels.py:225-244 (per-point loop)."""
    c, rg, tg, pg, p = coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points)
    out = np.zeros((p.shape[0], 11))
    for i, pt in enumerate(p):
        out[i] = sample_one(c, rg, tg, pg, pt)
    return out


def stokes_rk2_brick(coeffs, ds_cm, initial=None) -> np.ndarray:
    """native_kernels.py:279-294 == lib.rs:210-225: k1 = f(S); k2 = f(S + ds/2 k1); S + ds k2."""
    c, prefix = coerce_coefficients(coeffs)
    ds = validate_ds(ds_cm)
    init = coerce_initial(initial, prefix)
    cf = c.reshape((-1, 11))
    s = init.reshape((-1, 4))
    k1 = stokes_rhs(s, cf)
    mid = s + 0.5 * ds * k1
    k2 = stokes_rhs(mid, cf)
    return np.ascontiguousarray((s + ds * k2).reshape(tuple(prefix) + (4,)))


def sample_and_step_stokes(coeffs, r_grid, theta_grid, phi_grid, points, ds_cm, initial=None,
                           invalid_policy="nan") -> np.ndarray:
    """native_kernels.py:325-351: sample, RK2 on valid rows, policy on invalid rows."""
    c, rg, tg, pg, p = coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points)
    ds = validate_ds(ds_cm)
    policy = coerce_policy(invalid_policy)
    init = coerce_initial(initial, (p.shape[0],))
    sampled = sample_brick_trilinear(c, rg, tg, pg, p)
    out = np.full((p.shape[0], 4), np.nan)
    valid = sample_brick_valid_mask(rg, tg, p)
    if np.any(valid):
        out[valid] = stokes_rk2_brick(sampled[valid], ds, init[valid])
    if np.any(~valid):
        if policy == "initial":
            out[~valid] = init[~valid]
        elif policy == "zero":
            out[~valid] = 0.0
    return np.ascontiguousarray(out)


# --- vectorised equivalents -------------------------------------------------

def wrap_phi_vec(phi: np.ndarray, base: float) -> np.ndarray:
    """np.mod is the same floor modulo as Python's float %."""
    return np.mod(phi - base, TWO_PI) + base


def bracket_linear_vec(grid: np.ndarray, x: np.ndarray):
    """Vectorised bracket_linear; returns (i0, w, valid).  NaN x is 'valid' with NaN weight."""
    n = grid.size
    outside = (x < grid[0]) | (x > grid[-1])
    i0 = np.searchsorted(grid, x, side="right") - 1
    i0 = np.clip(i0, 0, n - 2)
    w = (x - grid[i0]) / np.maximum(grid[i0 + 1] - grid[i0], 1.0e-300)
    last = x == grid[-1]
    i0 = np.where(last, n - 2, i0)
    w = np.where(last, 1.0, w)
    return i0, w, ~outside


def bracket_phi_vec(grid: np.ndarray, phi: np.ndarray):
    n = grid.size
    p = wrap_phi_vec(phi, float(grid[0]))
    i0 = np.searchsorted(grid, p, side="right") - 1
    i0 = np.where(i0 < 0, n - 1, i0)
    i0 = np.where(i0 >= n, n - 1, i0)
    i1 = (i0 + 1) % n
    hi = np.where(i1 > i0, grid[i1], grid[0] + TWO_PI)
    pp = np.where(p >= grid[i0], p, p + TWO_PI)
    w = (pp - grid[i0]) / np.maximum(hi - grid[i0], 1.0e-300)
    return i0, i1, w


def sample_vec(coeffs: np.ndarray, rg, tg, pg, r, th, ph):
    """Vectorised sample_one; returns ((n, 11) coefficients, valid mask).  Invalid rows are NaN."""
    r0, wr, vr = bracket_linear_vec(rg, r)
    t0, wt, vt = bracket_linear_vec(tg, th)
    p0, p1, wp = bracket_phi_vec(pg, ph)
    r1, t1 = r0 + 1, t0 + 1
    wp_, wt_, wr_ = wp[:, None], wt[:, None], wr[:, None]
    c00 = (1.0 - wp_) * coeffs[r0, t0, p0] + wp_ * coeffs[r0, t0, p1]
    c01 = (1.0 - wp_) * coeffs[r0, t1, p0] + wp_ * coeffs[r0, t1, p1]
    c10 = (1.0 - wp_) * coeffs[r1, t0, p0] + wp_ * coeffs[r1, t0, p1]
    c11 = (1.0 - wp_) * coeffs[r1, t1, p0] + wp_ * coeffs[r1, t1, p1]
    c0 = (1.0 - wt_) * c00 + wt_ * c01
    c1 = (1.0 - wt_) * c10 + wt_ * c11
    out = (1.0 - wr_) * c0 + wr_ * c1
    valid = vr & vt
    out[~valid] = np.nan
    return out, valid


def sample_vec_numpy32(coeffs32: np.ndarray, rg32, tg32, pg32, r, th, ph):
    """The reference renderer's sample_coefficient_brick (accelerated_renderer.py:73-97, brackets of
    grmhd.py:229-257) on a float32 brick, as NumPy evaluates it.

    There the coordinates are Python floats and the grids and coefficient arrays float32; NumPy casts a
    Python float that meets a float32 scalar or array to float32, so the range tests, the searchsorted
    brackets, the weights (x - g0) / (g1 - g0) and the three lerp levels all run in float32 arithmetic.
    The phi wrap and the seam shift p + 2 pi stay Python (double) arithmetic; grid[0] + 2 pi is a
    float32 sum.  Returns ((n, 11) float64 coefficients, valid mask); invalid rows are NaN.
    """
    f32 = np.float32
    rg, tg, pg = (np.asarray(g, dtype=f32) for g in (rg32, tg32, pg32))
    c = np.asarray(coeffs32, dtype=f32)
    with np.errstate(invalid="ignore"):
        xf, yf = np.asarray(r, dtype=np.float64).astype(f32), np.asarray(th, dtype=np.float64).astype(f32)
        valid = ~((xf < rg[0]) | (xf > rg[-1]) | (yf < tg[0]) | (yf > tg[-1]))

        def bracket(g, x):
            i0 = np.clip(np.searchsorted(g, x, side="right") - 1, 0, g.size - 2)
            w = ((x - g[i0]) / (g[i0 + 1] - g[i0])).astype(f32)
            return i0, np.where(x == g[-1], f32(1.0), w).astype(f32)

        r0, wr = bracket(rg, xf)
        t0, wt = bracket(tg, yf)
        p = wrap_phi_vec(np.asarray(ph, dtype=np.float64), float(pg[0]))
        pf = p.astype(f32)
        n = pg.size
        p0 = np.searchsorted(pg, pf, side="right") - 1
        p0 = np.where(p0 < 0, n - 1, np.minimum(p0, n - 1))
        p1 = (p0 + 1) % n
        hi = np.where(p1 > p0, pg[p1], f32(pg[0] + f32(TWO_PI)))
        pp = np.where(pf >= pg[p0], p, p + TWO_PI)
        wp = ((pp.astype(f32) - pg[p0]) / (hi - pg[p0])).astype(f32)
        r1, t1 = r0 + 1, t0 + 1
        ap, at, ar = (f32(1.0) - wp)[:, None], (f32(1.0) - wt)[:, None], (f32(1.0) - wr)[:, None]
        wp_, wt_, wr_ = wp[:, None], wt[:, None], wr[:, None]
        c00 = ap * c[r0, t0, p0] + wp_ * c[r0, t0, p1]
        c01 = ap * c[r0, t1, p0] + wp_ * c[r0, t1, p1]
        c10 = ap * c[r1, t0, p0] + wp_ * c[r1, t0, p1]
        c11 = ap * c[r1, t1, p0] + wp_ * c[r1, t1, p1]
        c0 = at * c00 + wt_ * c01
        c1 = at * c10 + wt_ * c11
        out = (ar * c0 + wr_ * c1).astype(np.float64)
    out[~valid] = np.nan
    return out, valid


def sample_brick_trilinear_vec(coeffs, r_grid, theta_grid, phi_grid, points) -> np.ndarray:
    c, rg, tg, pg, p = coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points)
    out, _ = sample_vec(c, rg, tg, pg, p[:, 0], p[:, 1], p[:, 2])
    return out


def sample_and_step_stokes_vec(coeffs, r_grid, theta_grid, phi_grid, points, ds_cm, initial=None,
                               invalid_policy="nan") -> np.ndarray:
    c, rg, tg, pg, p = coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points)
    ds = validate_ds(ds_cm)
    policy = coerce_policy(invalid_policy)
    init = coerce_initial(initial, (p.shape[0],))
    sampled, valid = sample_vec(c, rg, tg, pg, p[:, 0], p[:, 1], p[:, 2])
    out = np.full((p.shape[0], 4), np.nan)
    if np.any(valid):
        k1 = stokes_rhs(init[valid], sampled[valid])
        k2 = stokes_rhs(init[valid] + 0.5 * ds * k1, sampled[valid])
        out[valid] = init[valid] + ds * k2
    if policy == "initial":
        out[~valid] = init[~valid]
    elif policy == "zero":
        out[~valid] = 0.0
    return out
