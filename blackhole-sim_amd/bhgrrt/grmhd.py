"""GRMHD snapshot fields the coefficient precompute reads, and the reference's analytic torus fixture.

``GRMHDSnapshot`` mirrors python/blackhole_sim/grmhd.py:55-95 (grids, rho, theta_e, pressure, b_con,
spin) with the same schema checks; ``u_con`` is optional here because the coefficient precompute
(coefficient_bricks.py:93-152) never reads it.  ``generate_analytic_grmhd_torus`` restates
grmhd.py:320-391 for the fields the precompute reads (rho, theta_e, pressure, b_con); it does not
build the four-velocity.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .scenes import horizon_radius


class SnapshotSchemaError(ValueError):
    """grmhd.py:30-31."""


def isco_radius(a: float, prograde: bool = True) -> float:
    """Equatorial ISCO radius (kerr.py:387-394, Bardeen, Press and Teukolsky 1972)."""
    if not abs(a) < 1.0:
        raise ValueError("spin must satisfy |a| < 1")
    aa = abs(a)
    z1 = 1.0 + (1.0 - aa * aa) ** (1.0 / 3.0) * ((1.0 + aa) ** (1.0 / 3.0) + (1.0 - aa) ** (1.0 / 3.0))
    z2 = math.sqrt(3.0 * aa * aa + z1 * z1)
    sign = -1.0 if prograde else 1.0
    return 3.0 + z2 + sign * math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))


@dataclass(frozen=True)
class GRMHDSnapshot:
    """grmhd.py:55-95."""

    r: np.ndarray
    theta: np.ndarray
    phi: np.ndarray
    rho: np.ndarray
    theta_e: np.ndarray
    pressure: np.ndarray
    b_con: np.ndarray
    spin_a: float
    u_con: np.ndarray | None = None
    time_m: float = 0.0
    metadata: dict[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        for name in ("r", "theta", "phi"):
            arr = np.asarray(getattr(self, name), dtype=float)
            if arr.ndim != 1 or arr.size < 2 or not np.all(np.diff(arr) > 0):
                raise SnapshotSchemaError(f"{name} grid must be a strictly increasing 1-D array")
            object.__setattr__(self, name, arr)
        shape = self.shape
        for name in ("rho", "theta_e", "pressure"):
            arr = np.asarray(getattr(self, name), dtype=float)
            if arr.shape != shape:
                raise SnapshotSchemaError(f"{name} must have shape {shape}; got {arr.shape}")
            object.__setattr__(self, name, arr)
        for name in ("b_con", "u_con"):
            val = getattr(self, name)
            if val is None:
                continue
            arr = np.asarray(val, dtype=float)
            if arr.shape != shape + (4,):
                raise SnapshotSchemaError(f"{name} must have shape {shape + (4,)}; got {arr.shape}")
            object.__setattr__(self, name, arr)
        if self.b_con is None:
            raise SnapshotSchemaError("b_con is required")
        if abs(float(self.spin_a)) >= 1.0:
            raise SnapshotSchemaError("spin_a must satisfy |a| < 1")

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.r.size, self.theta.size, self.phi.size)


def generate_analytic_grmhd_torus(spin_a: float = 0.85, nr: int = 72, ntheta: int = 40, nphi: int = 48,
                                  r_min: float | None = None, r_max: float = 60.0,
                                  density_peak_radius: float | None = None, scale_height: float = 0.32,
                                  electron_temp_peak: float = 18.0, magnetization: float = 0.08,
                                  time_m: float = 0.0) -> GRMHDSnapshot:
    """The deterministic torus of grmhd.py:320-391: Gaussian in log r and theta with a two-armed spiral
    and a small ripple, power-law electron temperature, toroidal field with small poloidal parts."""
    r_plus = horizon_radius(spin_a)
    rin = max(isco_radius(spin_a, prograde=True), r_plus * 1.15)
    if r_min is None:
        r_min = max(r_plus * 1.08, rin * 0.75)
    r0 = density_peak_radius if density_peak_radius is not None else max(8.0, 2.7 * rin)
    r = np.geomspace(r_min, r_max, nr)
    theta = np.linspace(0.045, math.pi - 0.045, ntheta)
    phi = np.linspace(0.0, 2.0 * math.pi, nphi, endpoint=False)
    R, TH, PH = np.meshgrid(r, theta, phi, indexing="ij")

    log_width = 0.46
    vertical = (TH - math.pi / 2.0) / scale_height
    radial = np.log(R / r0) / log_width
    spiral = 1.0 + 0.08 * np.sin(2.0 * PH + 0.7 * np.log(R) - 0.03 * time_m)
    ripple = 1.0 + 0.035 * np.sin(5.0 * PH + 3.0 * TH + 0.17 * R)
    rho = np.maximum(np.exp(-0.5 * (radial * radial + vertical * vertical)) * spiral * ripple, 1.0e-8)

    theta_e = electron_temp_peak * (np.maximum(R, rin) / rin) ** -0.84 * (0.82 + 0.18 * np.exp(-0.5 * vertical * vertical))
    theta_e = np.maximum(theta_e, 1.0e-3)
    pressure = rho * theta_e

    b_con = np.zeros(R.shape + (4,), dtype=float)
    b_strength = magnetization * np.sqrt(np.maximum(rho, 0.0)) / np.maximum(R, 1.0)
    b_con[..., 3] = b_strength / np.maximum(np.sin(TH), 1.0e-3)
    b_con[..., 1] = 0.04 * b_strength * np.sin(PH)
    b_con[..., 2] = 0.03 * b_strength * np.cos(PH)
    return GRMHDSnapshot(r=r, theta=theta, phi=phi, rho=rho, theta_e=theta_e, pressure=pressure, b_con=b_con,
                         spin_a=spin_a, time_m=time_m,
                         metadata={"generator": "analytic_torus_fixture", "not_grmhd_solver_output": True})


__all__ = ["GRMHDSnapshot", "SnapshotSchemaError", "generate_analytic_grmhd_torus", "isco_radius"]
