"""Drop-in for the reference's coefficient-brick precompute (python/blackhole_sim/coefficient_bricks.py:118-152).

``precompute_coefficient_bricks(snapshot, scaling, nu_hz, dtype, coeff_model, stride)`` has the
reference's signature and result (a ``CoefficientBrickGrid`` with the snapshot grids, the (nr, nt, np, 11)
coefficients in ``dtype``, ``nu_hz`` and ``spin_a``); the coefficients come from one FP64 HIP kernel
(one lane per cell) instead of the per-cell Python loop.  ``coeff_model`` is the reference's
``HybridSynchrotronCoefficients`` (synchrotron.py:315-325) with its thermal and power-law parameters;
``PhysicalScaling`` and ``BlackHoleSystem`` mirror calibration.py:28-119 and physics.py:22-64 for the
fields the precompute reads.  There is no CPU path: without the library this module raises.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Literal

import numpy as np

from . import _lib
from .scenes import Brick, COEFF_NAMES

# physics.py:14-19, calibration.py:18-26
G_SI = 6.67430e-11
C_SI = 299_792_458.0
M_SUN_SI = 1.98847e30
PARSEC_SI = 3.085677581491367e16
KPC_SI = 1.0e3 * PARSEC_SI
MPC_SI = 1.0e6 * PARSEC_SI
C_CGS = 2.99792458e10
G_CGS = 6.67430e-8
M_SUN_CGS = 1.98847e33


@dataclass(frozen=True)
class BlackHoleSystem:
    """physics.py:22-51 (mass, spin, distance)."""

    mass_kg: float
    spin_a: float = 0.0
    distance_m: float = 1.0 * KPC_SI

    @classmethod
    def sgr_a_star(cls, spin_a: float = 0.0) -> "BlackHoleSystem":
        return cls(mass_kg=4.297e6 * M_SUN_SI, spin_a=spin_a, distance_m=8.277 * KPC_SI)

    @classmethod
    def m87_star(cls, spin_a: float = 0.5) -> "BlackHoleSystem":
        return cls(mass_kg=6.5e9 * M_SUN_SI, spin_a=spin_a, distance_m=16.8 * MPC_SI)


@dataclass(frozen=True)
class PhysicalScaling:
    """calibration.py:28-119: code units -> cgs."""

    mass_bh_g: float
    distance_cm: float
    rho_cgs_per_code: float
    b_gauss_per_code: float
    ion_to_electron_temperature_ratio: float = 3.0
    mean_molecular_weight_per_electron: float = 1.0
    flux_scale_jy: float = 1.0

    @classmethod
    def from_black_hole_system(cls, system: BlackHoleSystem, rho_cgs_per_code: float, b_gauss_per_code: float,
                               **kwargs) -> "PhysicalScaling":
        return cls(mass_bh_g=system.mass_kg * 1.0e3, distance_cm=system.distance_m * 100.0,
                   rho_cgs_per_code=float(rho_cgs_per_code), b_gauss_per_code=float(b_gauss_per_code), **kwargs)

    @classmethod
    def from_mdot(cls, system: BlackHoleSystem, mdot_msun_per_year: float, density_geometry_factor: float = 4.0 * math.pi,
                  b_plasma_beta: float = 10.0, pressure_over_rho_c2: float = 1.0e-3, **kwargs) -> "PhysicalScaling":
        """calibration.py:72-101: density unit mdot / (geometry r_g^2 c), B from the plasma beta."""
        mass_g = system.mass_kg * 1.0e3
        rg_cm = G_CGS * mass_g / (C_CGS * C_CGS)
        mdot_g_s = mdot_msun_per_year * M_SUN_CGS / (365.25 * 24.0 * 3600.0)
        rho_unit = mdot_g_s / max(density_geometry_factor * rg_cm * rg_cm * C_CGS, 1.0e-300)
        p_gas = pressure_over_rho_c2 * rho_unit * C_CGS * C_CGS
        b_unit = math.sqrt(max(8.0 * math.pi * p_gas / max(b_plasma_beta, 1.0e-300), 0.0))
        return cls(mass_bh_g=mass_g, distance_cm=system.distance_m * 100.0, rho_cgs_per_code=rho_unit,
                   b_gauss_per_code=b_unit, **kwargs)


@dataclass(frozen=True)
class ThermalSynchrotronCoefficients:
    """synchrotron.py:231-238 (parameters)."""

    gamma_grid_size: int = 192
    pitch_nodes: int = 16
    faraday_rotation_scale: float = 1.0
    faraday_conversion_scale: float = 1.0


@dataclass(frozen=True)
class NonthermalPowerLawSynchrotronCoefficients:
    """synchrotron.py:271-282 (parameters)."""

    p: float = 3.5
    gamma_min: float = 10.0
    gamma_max: float = 1.0e5
    nonthermal_fraction: float = 0.02
    gamma_grid_size: int = 224
    pitch_nodes: int = 16
    absorption_scale: float = 1.0
    faraday_scale: float = 0.15


@dataclass(frozen=True)
class HybridSynchrotronCoefficients:
    """synchrotron.py:315-319 (thermal + power law)."""

    thermal: ThermalSynchrotronCoefficients = field(default_factory=ThermalSynchrotronCoefficients)
    nonthermal: NonthermalPowerLawSynchrotronCoefficients = field(default_factory=NonthermalPowerLawSynchrotronCoefficients)


@dataclass(frozen=True)
class CoefficientBrickGrid:
    """coefficient_bricks.py:28-75: grids, (nr, nt, np, 11) coefficients in ``dtype_name``, nu and spin."""

    r: np.ndarray
    theta: np.ndarray
    phi: np.ndarray
    coeffs: np.ndarray
    nu_hz: float
    spin_a: float
    dtype_name: str = "float32"

    @property
    def shape(self) -> tuple[int, int, int, int]:
        return tuple(self.coeffs.shape)

    @property
    def bytes(self) -> int:
        return int(self.coeffs.nbytes + 4 * (self.r.size + self.theta.size + self.phi.size))

    def to_brick(self) -> Brick:
        """The renderer's brick (float32 grids and coefficients)."""
        return Brick(self.r, self.theta, self.phi, self.coeffs.astype(np.float32),
                     meta={"nu_hz": self.nu_hz, "spin_a": self.spin_a})

    def save_npz(self, path) -> None:
        """coefficient_bricks.py:57-70."""
        np.savez_compressed(path, r=self.r.astype(np.float32), theta=self.theta.astype(np.float32),
                            phi=self.phi.astype(np.float32), coeffs=self.coeffs, coeff_names=np.array(COEFF_NAMES),
                            nu_hz=np.array(self.nu_hz), spin_a=np.array(self.spin_a),
                            dtype_name=np.array(self.dtype_name))


def synchrotron_desc(scaling: PhysicalScaling, nu_hz: float,
                     coeff_model: HybridSynchrotronCoefficients | None = None) -> _lib.SynchrotronDesc:
    m = coeff_model or HybridSynchrotronCoefficients()
    th, nt = m.thermal, m.nonthermal
    return _lib.SynchrotronDesc(
        nu_hz=float(nu_hz), rho_cgs_per_code=float(scaling.rho_cgs_per_code),
        b_gauss_per_code=float(scaling.b_gauss_per_code),
        mean_molecular_weight_per_electron=float(scaling.mean_molecular_weight_per_electron),
        th_gamma_grid_size=int(th.gamma_grid_size), th_pitch_nodes=int(th.pitch_nodes),
        th_faraday_rotation_scale=float(th.faraday_rotation_scale),
        th_faraday_conversion_scale=float(th.faraday_conversion_scale),
        nt_p=float(nt.p), nt_gamma_min=float(nt.gamma_min), nt_gamma_max=float(nt.gamma_max),
        nt_fraction=float(nt.nonthermal_fraction), nt_gamma_grid_size=int(nt.gamma_grid_size),
        nt_pitch_nodes=int(nt.pitch_nodes), nt_absorption_scale=float(nt.absorption_scale),
        nt_faraday_scale=float(nt.faraday_scale))


def precompute_coefficient_bricks(snapshot, scaling: PhysicalScaling, nu_hz: float = 230e9,
                                  dtype: Literal["float32", "float16", "float64"] = "float32",
                                  coeff_model: HybridSynchrotronCoefficients | None = None,
                                  stride: int = 1) -> CoefficientBrickGrid:
    """coefficient_bricks.py:118-152 on the GPU.  ``snapshot`` needs r, theta, phi, rho, theta_e, b_con and
    spin_a (a ``GRMHDSnapshot`` of this package or of the reference)."""
    if stride < 1:
        raise ValueError("stride must be >= 1")
    if dtype not in _lib.BRICK_DTYPE:
        raise ValueError("dtype must be float32, float16, or float64")
    sl = (slice(None, None, stride),) * 3
    rho = np.ascontiguousarray(np.asarray(snapshot.rho, dtype=np.float64)[sl])
    theta_e = np.ascontiguousarray(np.asarray(snapshot.theta_e, dtype=np.float64)[sl])
    b_con = np.ascontiguousarray(np.asarray(snapshot.b_con, dtype=np.float64)[sl])
    shape = rho.shape
    if theta_e.shape != shape or b_con.shape != shape + (4,):
        raise ValueError("rho, theta_e and b_con must cover the same grid (b_con with 4 components)")
    out = np.empty(shape + (11,), dtype=np.dtype(dtype))
    desc = synchrotron_desc(scaling, nu_hz, coeff_model)
    _lib.check(_lib.lib().bh_precompute_coefficient_bricks(_lib.dptr(rho), _lib.dptr(theta_e), _lib.dptr(b_con),
                                                           C.c_int64(rho.size), C.byref(desc),
                                                           _lib.BRICK_DTYPE[dtype], out.ctypes.data_as(C.c_void_p)))
    grids = [np.asarray(g, dtype=np.float64)[::stride] for g in (snapshot.r, snapshot.theta, snapshot.phi)]
    return CoefficientBrickGrid(*grids, out, float(nu_hz), float(snapshot.spin_a), dtype)


def estimate_brick_memory(nr: int, ntheta: int, nphi: int, dtype: str = "float32") -> dict[str, float]:
    """coefficient_bricks.py:155-160."""
    coeff_bytes = int(nr) * int(ntheta) * int(nphi) * len(COEFF_NAMES) * np.dtype(dtype).itemsize
    grid_bytes = (int(nr) + int(ntheta) + int(nphi)) * 4
    total = coeff_bytes + grid_bytes
    return {"coeff_bytes": float(coeff_bytes), "grid_bytes": float(grid_bytes), "total_bytes": float(total),
            "total_mib": total / (1024.0 * 1024.0)}


__all__ = ["BlackHoleSystem", "PhysicalScaling", "ThermalSynchrotronCoefficients",
           "NonthermalPowerLawSynchrotronCoefficients", "HybridSynchrotronCoefficients", "CoefficientBrickGrid",
           "synchrotron_desc", "precompute_coefficient_bricks", "estimate_brick_memory"]
