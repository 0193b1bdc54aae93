"""Drop-in replacement for the reference's ``blackhole_native`` extension module.

The reference loads ``importlib.import_module("blackhole_native")``
(python/blackhole_sim/native_loader.py:13-17) and calls the five PyO3
functions of native/core/src/lib.rs:368-376 with flat float lists.  Putting
this package's parent directory on ``sys.path`` makes the reference's own
``native_kernels.py`` run on the MI355X through the C ABI.

Inputs may be lists or float64 arrays (arrays are passed zero-copy); results
are flat float64 NumPy arrays, which ``np.asarray(...).reshape(...)`` in the
reference wrapper consumes unchanged.  Errors raise ValueError with the Rust
messages.
"""

from __future__ import annotations

import numpy as np

from bhgrrt import _lib

__all__ = ["core_version", "detect_arch", "stokes_rk2_brick", "sample_brick_trilinear", "sample_and_step_stokes",
           "sample_brick_valid_mask"]


def _flat(values) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())


def core_version() -> str:
    """lib.rs:19-22."""
    return _lib.lib().bh_core_version().decode()


def detect_arch() -> dict:
    """lib.rs:24-35: {arch, os, family, pointer_width} as strings."""
    import ctypes

    buf = ctypes.create_string_buffer(256)
    _lib.check(_lib.lib().bh_detect_arch(buf, 256))
    return dict(kv.split("=", 1) for kv in buf.value.decode().split(";"))


def stokes_rk2_brick(coeffs, ds_cm: float, initial) -> np.ndarray:
    """lib.rs:227-279."""
    c, i = _flat(coeffs), _flat(initial)
    cells = c.size // 11 if c.size % 11 == 0 else 0
    out = np.empty(max(cells, 1) * 4)
    _lib.check(_lib.lib().bh_stokes_rk2_brick(_lib.dptr(c), c.size, float(ds_cm), _lib.dptr(i), i.size,
                                              _lib.dptr(out)))
    return out[: cells * 4]


def sample_brick_trilinear(coeffs, r_grid, theta_grid, phi_grid, points) -> np.ndarray:
    """lib.rs:281-306."""
    c, rg, tg, pg, p = map(_flat, (coeffs, r_grid, theta_grid, phi_grid, points))
    n = p.size // 3
    out = np.empty(max(n, 1) * 11)
    _lib.check(_lib.lib().bh_sample_brick_trilinear(_lib.dptr(c), c.size, _lib.dptr(rg), rg.size, _lib.dptr(tg),
                                                    tg.size, _lib.dptr(pg), pg.size, _lib.dptr(p), p.size,
                                                    _lib.dptr(out)))
    return out[: n * 11]


def sample_and_step_stokes(coeffs, r_grid, theta_grid, phi_grid, points, ds_cm: float, initial) -> np.ndarray:
    """lib.rs:308-366: invalid points give NaN rows (policy 'nan')."""
    c, rg, tg, pg, p, i = map(_flat, (coeffs, r_grid, theta_grid, phi_grid, points, initial))
    n = p.size // 3
    out = np.empty(max(n, 1) * 4)
    _lib.check(_lib.lib().bh_sample_and_step_stokes(_lib.dptr(c), c.size, _lib.dptr(rg), rg.size, _lib.dptr(tg),
                                                    tg.size, _lib.dptr(pg), pg.size, _lib.dptr(p), p.size,
                                                    float(ds_cm), _lib.dptr(i), i.size, _lib.POLICY["nan"],
                                                    _lib.dptr(out)))
    return out[: n * 4]


def sample_brick_valid_mask(r_grid, theta_grid, points) -> np.ndarray:
    """native_kernels.py:189-210 on the GPU: bool mask of in-domain r/theta."""
    rg, tg, p = map(_flat, (r_grid, theta_grid, points))
    n = p.size // 3
    out = np.zeros(max(n, 1), dtype=np.uint8)
    _lib.check(_lib.lib().bh_sample_brick_valid_mask(_lib.dptr(rg), rg.size, _lib.dptr(tg), tg.size, _lib.dptr(p),
                                                     p.size, out.ctypes.data_as(_lib._u8p)))
    return out[:n].astype(bool)
