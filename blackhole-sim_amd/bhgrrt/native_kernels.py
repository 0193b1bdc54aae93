"""Host-side mirror of python/blackhole_sim/native_kernels.py backed by the HIP library.

Same public names, argument meaning, shapes, validation messages and
``invalid_policy`` semantics as the reference module; the computation runs in
FP64 on the GPU through the C ABI (include/bhgrrt.h).  ``prefer_native`` is
accepted for signature compatibility: this build has no CPU path, so the HIP
kernels always run (and fail loudly without a GPU).
"""

from __future__ import annotations

from typing import Any, Literal

import numpy as np

from . import _lib

STOKES_RK2_RTOL = 1.0e-12          # native_kernels.py:12-15
STOKES_RK2_ATOL = 1.0e-12
SAMPLE_BRICK_TRILINEAR_RTOL = 1.0e-12
SAMPLE_BRICK_TRILINEAR_ATOL = 1.0e-12
InvalidSamplePolicy = Literal["nan", "initial", "zero"]


def deterministic_stokes_coefficients(nr: int = 4, ntheta: int = 3, nphi: int = 4,
                                      dtype: str | np.dtype[Any] = "float64") -> np.ndarray:
    """Deterministic parity brick with the formulas of native_kernels.py:19-50."""
    if min(int(nr), int(ntheta), int(nphi)) < 1:
        raise ValueError("nr, ntheta, and nphi must be positive")
    r = np.linspace(0.0, 1.0, int(nr))
    t = np.linspace(0.0, 1.0, int(ntheta))
    p = np.linspace(0.0, 1.0, int(nphi))
    rr, tt, pp = np.meshgrid(r, t, p, indexing="ij")
    c = np.empty((int(nr), int(ntheta), int(nphi), 11))
    c[..., 0] = 1.0e-3 * (1.0 + rr)
    c[..., 1] = 2.0e-4 * (tt - 0.5)
    c[..., 2] = 1.5e-4 * (pp - 0.5)
    c[..., 3] = 1.0e-4 * (rr - tt)
    c[..., 4] = 1.0e-2 + 1.0e-3 * rr
    c[..., 5] = 1.0e-4 * (tt - 0.5)
    c[..., 6] = 1.0e-4 * (pp - 0.5)
    c[..., 7] = 8.0e-5 * (rr - 0.5)
    c[..., 8] = 2.0e-3 * (pp - 0.5)
    c[..., 9] = 1.0e-3 * (rr - tt)
    c[..., 10] = 1.0e-3 * (tt - pp)
    return c.astype(dtype, copy=False)


# --- validation, as native_kernels.py:53-130 --------------------------------

def _coerce_coefficients(coeffs):
    arr = np.asarray(coeffs, dtype=np.float64)
    if arr.ndim < 1 or arr.shape[-1] != 11:
        raise ValueError("coeffs must have shape (..., 11)")
    if arr.size == 0:
        raise ValueError("coeffs must contain at least one cell")
    if not np.all(np.isfinite(arr)):
        raise ValueError("coeffs must contain only finite values")
    return np.ascontiguousarray(arr), tuple(int(v) for v in arr.shape[:-1])


def _coerce_grid(name, grid):
    g = np.asarray(grid, dtype=np.float64)
    if g.ndim != 1 or g.size < 2:
        raise ValueError(f"{name} grid must contain at least 2 values")
    if not np.all(np.isfinite(g)):
        raise ValueError(f"{name} grid must contain only finite values")
    if np.any(np.diff(g) <= 0.0):
        raise ValueError(f"{name} grid must be strictly increasing")
    return np.ascontiguousarray(g)


def _coerce_points(points):
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


def _coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points):
    c, prefix = _coerce_coefficients(coeffs)
    if len(prefix) != 3:
        raise ValueError("coeffs must have shape (r, theta, phi, 11)")
    rg = _coerce_grid("r", r_grid)
    tg = _coerce_grid("theta", theta_grid)
    pg = _coerce_grid("phi", phi_grid)
    expect = (int(rg.size), int(tg.size), int(pg.size))
    if prefix != expect:
        raise ValueError(f"coeffs grid shape must be {expect}; got {prefix}")
    return c, rg, tg, pg, _coerce_points(points)


def _coerce_initial(initial, prefix):
    if initial is None:
        return np.zeros(tuple(prefix) + (4,))
    arr = np.asarray(initial, dtype=np.float64)
    if not np.all(np.isfinite(arr)):
        raise ValueError("initial must contain only finite values")
    if arr.shape == (4,):
        return np.ascontiguousarray(np.broadcast_to(arr, tuple(prefix) + (4,)))
    if arr.shape != tuple(prefix) + (4,):
        raise ValueError(f"initial must have shape (4,) or {tuple(prefix) + (4,)}")
    return np.ascontiguousarray(arr)


def _validate_ds(ds_cm):
    v = float(ds_cm)
    if not np.isfinite(v) or v < 0.0:
        raise ValueError("ds_cm must be finite and non-negative")
    return v


def _coerce_invalid_sample_policy(policy):
    if policy not in _lib.POLICY:
        raise ValueError("invalid_policy must be 'nan', 'initial', or 'zero'")
    return policy


# --- public API ----------------------------------------------------------------

def native_sample_brick_trilinear_available() -> bool:
    return _available()


def native_stokes_rk2_available() -> bool:
    return _available()


def native_sample_and_step_stokes_available() -> bool:
    return _available()


def _available() -> bool:
    try:
        return _lib.device_count() > 0
    except ImportError:
        return False


def sample_brick_valid_mask(r_grid, theta_grid, points) -> np.ndarray:
    """native_kernels.py:189-210 on the GPU (bit-exact FP64 compares)."""
    rg = _coerce_grid("r", r_grid)
    tg = _coerce_grid("theta", theta_grid)
    p = _coerce_points(points)
    out = np.zeros(p.shape[0], dtype=np.uint8)
    _lib.check(_lib.lib().bh_sample_brick_valid_mask(_lib.dptr(rg), rg.size, _lib.dptr(tg), tg.size, _lib.dptr(p),
                                                     p.size, out.ctypes.data_as(_lib._u8p)))
    return out.astype(bool)


def sample_brick_trilinear(coeffs, r_grid, theta_grid, phi_grid, points, *, prefer_native: bool = True) -> np.ndarray:
    """native_kernels.py:252-276: (n, 11) FP64 samples, NaN rows outside r/theta."""
    c, rg, tg, pg, p = _coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points)
    out = np.empty((p.shape[0], 11))
    _lib.check(_lib.lib().bh_sample_brick_trilinear(_lib.dptr(c), c.size, _lib.dptr(rg), rg.size, _lib.dptr(tg),
                                                    tg.size, _lib.dptr(pg), pg.size, _lib.dptr(p), p.size,
                                                    _lib.dptr(out)))
    return out


def stokes_rk2_brick(coeffs, ds_cm: float, initial=None, *, prefer_native: bool = True) -> np.ndarray:
    """native_kernels.py:302-322: one RK2 Stokes step per cell."""
    c, prefix = _coerce_coefficients(coeffs)
    ds = _validate_ds(ds_cm)
    init = _coerce_initial(initial, prefix)
    cells = c.size // 11
    out = np.empty((cells, 4))
    _lib.check(_lib.lib().bh_stokes_rk2_brick(_lib.dptr(c), c.size, ds, _lib.dptr(init), init.size, _lib.dptr(out)))
    return out.reshape(tuple(prefix) + (4,))


def sample_and_step_stokes(coeffs, r_grid, theta_grid, phi_grid, points, ds_cm: float, initial=None, *,
                           prefer_native: bool = True, invalid_policy: InvalidSamplePolicy = "nan") -> np.ndarray:
    """native_kernels.py:359-407: sample + RK2 step, invalid rows per ``invalid_policy``."""
    c, rg, tg, pg, p = _coerce_sampler_inputs(coeffs, r_grid, theta_grid, phi_grid, points)
    ds = _validate_ds(ds_cm)
    policy = _coerce_invalid_sample_policy(invalid_policy)
    init = _coerce_initial(initial, (p.shape[0],))
    out = np.empty((p.shape[0], 4))
    _lib.check(_lib.lib().bh_sample_and_step_stokes(_lib.dptr(c), c.size, _lib.dptr(rg), rg.size, _lib.dptr(tg),
                                                    tg.size, _lib.dptr(pg), pg.size, _lib.dptr(p), p.size, ds,
                                                    _lib.dptr(init), init.size, _lib.POLICY[policy],
                                                    _lib.dptr(out)))
    return out
