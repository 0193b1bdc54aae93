"""Diagnostics for the HIP core, in the shape of the reference's doctor report.

``doctor_report()`` mirrors python/blackhole_sim/platform_probe.py:68-97 for the fields that concern
this library: whether the native core (``libbhgrrt.so``, exposed to the reference as
``blackhole_native``) loads, its version and architecture, the GPU devices it sees, and warnings.  The
reference's own ``doctor_report`` finds the same core through ``native_core_status``
(native_loader.py:31-61), which reads ``blackhole_native.core_version()`` and ``detect_arch()``.

``gpu_benchmark_rows()`` returns ``BenchmarkResult``-shaped rows (benchmark.py:32-45) for the GPU
kernels at the reference's benchmark sizes (benchmark.py:43-68, 86-104, 188-236), for a caller that
adds them to its ``blackhole_sim.benchmark.v2`` suite; the reference's parity benchmarks already time
these kernels as their "native" rows once ``blackhole_native`` imports.
"""

from __future__ import annotations

import ctypes as C
import platform
import time
from typing import Any

import numpy as np

from . import _lib

SCHEMA = "blackhole_sim.benchmark.v2"


def _devices() -> list[dict[str, Any]]:
    lib = _lib.lib()
    n = C.c_int(0)
    if lib.bh_device_count(C.byref(n)) != _lib.BH_OK:
        return []
    out = []
    for d in range(n.value):
        buf = C.create_string_buffer(256)
        cus = C.c_int(0)
        if lib.bh_device_name(d, buf, 256, C.byref(cus)) == _lib.BH_OK:
            out.append({"index": d, "name": buf.value.decode(), "compute_units": int(cus.value)})
    return out


def doctor_report() -> dict[str, Any]:
    """Native-core and GPU status (platform_probe.py:68-97 field names where they apply)."""
    warnings: list[str] = []
    try:
        lib = _lib.lib()
    except ImportError as exc:
        return {"native_core_loaded": False, "native_core_version": None, "native_core_arch": None,
                "gpu_backend": None, "devices": [], "library_path": str(_lib.LIB_PATH),
                "warnings": (f"native core not built: {exc}",)}
    from blackhole_native import detect_arch

    arch = detect_arch()
    devices = _devices()
    if not devices:
        warnings.append("no HIP device visible: the GPU entry points return BH_ERR_DEVICE")
    elif not any("gfx950" in d["name"] or "MI355" in d["name"] for d in devices):
        warnings.append("the kernels are built for gfx950 (MI355X) only")
    return {
        "process_arch": platform.machine().lower(),
        "native_core_loaded": True,
        "native_core_version": lib.bh_core_version().decode(),
        "native_core_arch": arch.get("arch"),
        "raw_arch": arch,
        "gpu_backend": "rocm" if devices else None,
        "devices": devices,
        "library_path": str(_lib.LIB_PATH),
        "warnings": tuple(warnings),
    }


def _row(name: str, seconds: float, iterations: int, items: int, **meta) -> dict[str, Any]:
    return {"name": name, "seconds": seconds, "iterations": iterations, "items": items,
            "items_per_second": items / seconds, "metadata": {"backend": "hip-gfx950", **meta}}


def gpu_benchmark_rows(nr: int = 8, ntheta: int = 6, nphi: int = 8, point_count: int = 64,
                       iterations: int = 1) -> list[dict[str, Any]]:
    """GPU rows for the reference's three benchmark workloads (host-buffer entry points, wall clock per
    call as the reference times them)."""
    if min(nr, ntheta, nphi, point_count, iterations) < 1:
        raise ValueError("nr, ntheta, nphi, point_count, and iterations must be positive")
    from . import coefficient_bricks as cb
    from . import native_kernels as nk
    from .grmhd import generate_analytic_grmhd_torus

    rows = []
    snap = generate_analytic_grmhd_torus(nr=nr, ntheta=ntheta, nphi=nphi)
    scaling = cb.PhysicalScaling(mass_bh_g=1.0, distance_cm=1.0, rho_cgs_per_code=1.0e-18, b_gauss_per_code=30.0)
    cb.precompute_coefficient_bricks(snap, scaling)
    t = time.perf_counter()
    for _ in range(iterations):
        cb.precompute_coefficient_bricks(snap, scaling)
    cells = nr * ntheta * nphi
    rows.append(_row("coefficient_brick_precompute", max(time.perf_counter() - t, 1e-12), iterations,
                     cells * iterations, grid=[nr, ntheta, nphi], dtype="float32"))

    coeffs = nk.deterministic_stokes_coefficients(nr=nr, ntheta=ntheta, nphi=nphi)
    r_grid = np.linspace(0.0, 1.0, nr)
    t_grid = np.linspace(0.0, 1.0, ntheta)
    p_grid = np.linspace(0.0, 2.0 * np.pi, nphi, endpoint=False)
    idx = np.arange(point_count, dtype=np.float64)
    den = max(point_count - 1, 1)
    pts = np.stack([0.05 + 0.90 * ((idx % den) / den), 0.10 + 0.80 * (((idx * 3.0) % (den + 1)) / den),
                    idx * 0.6180339887498949 * 2.0 * np.pi], axis=1)
    nk.sample_brick_trilinear(coeffs, r_grid, t_grid, p_grid, pts)
    t = time.perf_counter()
    for _ in range(iterations):
        nk.sample_brick_trilinear(coeffs, r_grid, t_grid, p_grid, pts)
    rows.append(_row("sample_brick_trilinear_native", max(time.perf_counter() - t, 1e-12), iterations,
                     point_count * iterations, grid=[nr, ntheta, nphi], point_count=point_count))

    initial = np.zeros(coeffs.shape[:-1] + (4,))
    initial[..., 0], initial[..., 1] = 1.0e-2, 1.0e-3
    nk.stokes_rk2_brick(coeffs, 0.05, initial)
    t = time.perf_counter()
    for _ in range(iterations):
        nk.stokes_rk2_brick(coeffs, 0.05, initial)
    rows.append(_row("stokes_rk2_brick_native", max(time.perf_counter() - t, 1e-12), iterations,
                     cells * iterations, grid=[nr, ntheta, nphi]))
    return rows


__all__ = ["SCHEMA", "doctor_report", "gpu_benchmark_rows"]
