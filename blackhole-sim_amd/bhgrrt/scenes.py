"""Scene, camera and coefficient-brick descriptions for the Kerr Stokes render.

The dataclasses mirror the reference's configuration objects:

* ``Camera`` follows ``LocalCamera`` (python/blackhole_sim/kerr.py:155-185):
  Boyer-Lindquist position of a ZAMO camera, vertical field of view and roll.
* ``Scene`` follows ``AcceleratedRenderConfig`` (python/blackhole_sim/
  accelerated_renderer.py:36-46) plus the transfer knobs of
  ``PolarizedTransferConfig`` (python/blackhole_sim/polarized_transfer.py:22-28)
  and the spin that ``GRMHDSnapshot.spin_a`` carries.  The step rule is either
  the fixed affine step of the reference tracer (kerr.py:337-346) or the
  adaptive rule h = h_frac * (r - r_+); the integrator is RK2 midpoint or the
  reference tracer's classical RK4 (kerr.py:306-313).
* ``Brick`` follows ``CoefficientBrickGrid`` (python/blackhole_sim/
  coefficient_bricks.py:28-75): float32 grids, AoS coefficients of shape
  (nr, ntheta, nphi, 11) in the channel order of ``COEFF_NAMES``
  (coefficient_bricks.py:21-25).

``make_brick`` draws the seeded synthetic brick used by every workload here:
j_I and alpha_I are r^-2 times independent lognormal(0, 0.5) draws, polarised
fractions are Uniform[0, 0.7] with a random EVPA, and rho ~ N(0, 0.1 alpha_I).
With ``disk=True`` the emissivities carry a thin equatorial profile
exp(-cos(theta)^2 / 0.02).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

COEFF_NAMES: tuple[str, ...] = (
    "j_i", "j_q", "j_u", "j_v",
    "alpha_i", "alpha_q", "alpha_u", "alpha_v",
    "rho_v", "rho_q", "rho_u",
)
N_COEFF = len(COEFF_NAMES)

STEP_FIXED = 0
STEP_ADAPTIVE = 1


def horizon_radius(a: float) -> float:
    """Outer horizon r_+ = 1 + sqrt(1 - a^2) (kerr.py:39-42)."""
    if not math.isfinite(a) or abs(a) >= 1.0:
        raise ValueError("Kerr spin must be finite and satisfy |a| < 1 for this integrator")
    return 1.0 + math.sqrt(1.0 - a * a)


@dataclass(frozen=True)
class Camera:
    """ZAMO camera; mirrors ``LocalCamera`` (kerr.py:155-185)."""

    r: float = 50.0
    theta: float = math.radians(60.0)
    phi: float = 0.0
    fov_y_degrees: float = 35.0
    roll_degrees: float = 0.0

    @classmethod
    def from_degrees(cls, r: float = 50.0, inclination_degrees: float = 60.0, phi_degrees: float = 0.0,
                     fov_y_degrees: float = 35.0, roll_degrees: float = 0.0) -> "Camera":
        return cls(r=r, theta=math.radians(inclination_degrees), phi=math.radians(phi_degrees),
                   fov_y_degrees=fov_y_degrees, roll_degrees=roll_degrees)

    @classmethod
    def spanning(cls, r: float, inclination_degrees: float, half_width_m: float, width: int, height: int,
                 phi_degrees: float = 0.0) -> "Camera":
        """Camera whose horizontal image half-width is ``half_width_m`` gravitational radii.

        The reference camera uses tan_x = aspect * tan(fov_y / 2) with
        aspect = width / height (kerr.py:209-211), so tan_x = half_width / r
        fixes the vertical field of view.
        """
        tan_x = half_width_m / r
        tan_y = tan_x / (width / height)
        return cls.from_degrees(r=r, inclination_degrees=inclination_degrees, phi_degrees=phi_degrees,
                                fov_y_degrees=math.degrees(2.0 * math.atan(tan_y)))


@dataclass(frozen=True)
class Scene:
    """One render: camera, spin, image size, step rule and transfer settings."""

    spin: float
    camera: Camera
    width: int
    height: int
    step_mode: int = STEP_FIXED
    step: float = 0.06            # fixed h, or h_frac for the adaptive rule
    max_steps: int = 5000
    escape_radius: float = 220.0
    affine_to_length_cm: float = 1.0
    max_optical_depth: float = 30.0
    exact_step: bool = False      # PolarizedTransferConfig.use_exact_matrix_step (else one RK2 step per segment)
    integrator: str = "rk2"       # geodesic step: "rk2" midpoint, or "rk4" as the reference tracer (kerr.py:306-313)
    # FP64 render: sample the float32 brick in float32 as the reference renderer does under NumPy's scalar
    # promotion (accelerated_renderer.py:73-97), instead of the drop-in sampler's FP64 (native_kernels.py:164-186)
    numpy32_sampling: bool = False

    @property
    def aspect(self) -> float:
        return self.width / self.height

    @property
    def rays(self) -> int:
        return self.width * self.height

    def with_size(self, width: int, height: int) -> "Scene":
        return replace(self, width=int(width), height=int(height))


@dataclass(frozen=True)
class Brick:
    """Coefficient brick; mirrors ``CoefficientBrickGrid`` (coefficient_bricks.py:28-55)."""

    r: np.ndarray
    theta: np.ndarray
    phi: np.ndarray
    coeffs: np.ndarray
    seed: int = -1
    meta: dict = field(default_factory=dict)

    def __post_init__(self) -> None:
        object.__setattr__(self, "r", np.ascontiguousarray(self.r, dtype=np.float32))
        object.__setattr__(self, "theta", np.ascontiguousarray(self.theta, dtype=np.float32))
        object.__setattr__(self, "phi", np.ascontiguousarray(self.phi, dtype=np.float32))
        c = np.ascontiguousarray(self.coeffs, dtype=np.float32)
        expect = (self.r.size, self.theta.size, self.phi.size, N_COEFF)
        if c.shape != expect:
            raise ValueError(f"coeffs must have shape {expect}; got {c.shape}")
        if not np.all(np.isfinite(c)):
            raise ValueError("coefficient brick contains non-finite values")
        object.__setattr__(self, "coeffs", c)

    @property
    def shape(self) -> tuple[int, int, int, int]:
        return self.coeffs.shape

    @property
    def nbytes(self) -> int:
        return int(self.coeffs.nbytes + self.r.nbytes + self.theta.nbytes + self.phi.nbytes)

    def save_npz(self, path) -> None:
        """Write the reference's npz schema (coefficient_bricks.py:57-70)."""
        np.savez_compressed(path, r=self.r, theta=self.theta, phi=self.phi, coeffs=self.coeffs,
                            coeff_names=np.array(COEFF_NAMES), nu_hz=np.array(self.meta.get("nu_hz", 230.0e9)),
                            spin_a=np.array(self.meta.get("spin_a", 0.0)),
                            dtype_name=np.array("float32"))

    @classmethod
    def load_npz(cls, path) -> "Brick":
        """Read the reference's npz schema (coefficient_bricks.py:72-75) without unpickling."""
        with np.load(path, allow_pickle=False) as z:
            return cls(z["r"], z["theta"], z["phi"], z["coeffs"],
                       meta={"nu_hz": float(z["nu_hz"]), "spin_a": float(z["spin_a"])})


def make_brick(nr: int, ntheta: int, nphi: int, spin: float, seed: int, r_max: float = 40.0,
               disk: bool = False) -> Brick:
    """Seeded synthetic coefficient brick over r in [1.05 r_+, r_max], theta in [0, pi], phi in [0, 2 pi).

    The r grid is geometric (as grmhd.py:345), theta is linspace [0, pi] and phi
    is linspace [0, 2 pi) without endpoint (as benchmark.py:97).
    """
    rp = horizon_radius(spin)
    r = np.geomspace(1.05 * rp, r_max, int(nr))
    theta = np.linspace(0.0, math.pi, int(ntheta))
    phi = np.linspace(0.0, 2.0 * math.pi, int(nphi), endpoint=False)
    rng = np.random.default_rng(int(seed))
    shape = (int(nr), int(ntheta), int(nphi))
    R = r[:, None, None] * np.ones(shape)
    TH = theta[None, :, None] * np.ones(shape)
    r2 = R ** -2.0
    j_i = r2 * rng.lognormal(0.0, 0.5, size=shape)
    alpha_i = r2 * rng.lognormal(0.0, 0.5, size=shape)
    if disk:
        j_i = j_i * np.exp(-np.cos(TH) ** 2 / 0.02)
    p_lin = rng.uniform(0.0, 0.7, size=shape)
    chi = rng.uniform(0.0, math.pi, size=shape)
    p_circ = rng.uniform(0.0, 0.7, size=shape) * np.where(rng.uniform(size=shape) < 0.5, -1.0, 1.0)
    c = np.empty(shape + (N_COEFF,), dtype=np.float64)
    c[..., 0] = j_i
    c[..., 1] = p_lin * j_i * np.cos(2.0 * chi)
    c[..., 2] = p_lin * j_i * np.sin(2.0 * chi)
    c[..., 3] = p_circ * j_i
    c[..., 4] = alpha_i
    c[..., 5] = p_lin * alpha_i * np.cos(2.0 * chi)
    c[..., 6] = p_lin * alpha_i * np.sin(2.0 * chi)
    c[..., 7] = p_circ * alpha_i
    c[..., 8] = rng.normal(0.0, 1.0, size=shape) * 0.1 * alpha_i   # rho_v
    c[..., 9] = rng.normal(0.0, 1.0, size=shape) * 0.1 * alpha_i   # rho_q
    c[..., 10] = rng.normal(0.0, 1.0, size=shape) * 0.1 * alpha_i  # rho_u
    return Brick(r, theta, phi, c.astype(np.float32), seed=int(seed),
                 meta={"spin_a": float(spin), "disk": bool(disk), "r_max": float(r_max)})


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

SMALL_FIXED_STEP = 20.0   # affine step that carries rays from r=1000 through the r<=40 brick in 64 steps


def small_plain_scene() -> Scene:
    """Plain method, small: a=0.9, r_obs=1000, i=60 deg, 32x18 pixels over +-12 M, 64 fixed RK2 steps."""
    cam = Camera.spanning(r=1000.0, inclination_degrees=60.0, half_width_m=12.0, width=32, height=18)
    return Scene(spin=0.9, camera=cam, width=32, height=18, step_mode=STEP_FIXED, step=SMALL_FIXED_STEP,
                 max_steps=64, escape_radius=220.0)


def small_brick() -> Brick:
    return make_brick(8, 6, 8, spin=0.9, seed=20240601, disk=False)


def scale_plain_scene(width: int = 1024, height: int = 1024) -> Scene:
    """Plain method at scale: a=0.94, i=17 deg, 1024^2 over +-20 M, h = 0.05 (r - r_+), <= 4096 steps."""
    cam = Camera.spanning(r=1000.0, inclination_degrees=17.0, half_width_m=20.0, width=width, height=height)
    return Scene(spin=0.94, camera=cam, width=width, height=height, step_mode=STEP_ADAPTIVE, step=0.05,
                 max_steps=4096, escape_radius=220.0)


def scale_brick(nr: int = 128, ntheta: int = 64, nphi: int = 128, seed: int = 20240602) -> Brick:
    return make_brick(nr, ntheta, nphi, spin=0.94, seed=seed, disk=True)


def pixel_ndc(scene: Scene, px: np.ndarray, py: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Pixel centre -> NDC as render_tile_stokes_bricks (accelerated_renderer.py:138-140)."""
    ndc_x = 2.0 * (np.asarray(px, dtype=np.float64) + 0.5) / scene.width - 1.0
    ndc_y = 1.0 - 2.0 * (np.asarray(py, dtype=np.float64) + 0.5) / scene.height
    return ndc_x, ndc_y


def batch_bricks(seed: int = 20240603) -> list[Brick]:
    """Four seeded disk bricks from 64x32x64 to 128x64x128 (grid from r_+ of a=0.998, so every spin fits)."""
    shapes = [(64, 32, 64), (80, 40, 80), (96, 48, 96), (128, 64, 128)]
    return [make_brick(nr, nt, np_, spin=0.998, seed=seed + k, disk=True) for k, (nr, nt, np_) in enumerate(shapes)]


def batch_scenes(n: int = 64, seed: int = 20240604, min_px: int = 128, max_px: int = 1024) -> list[tuple[Scene, int]]:
    """Batch workload: n renders, square sizes log-uniform in [min_px, max_px], a ~ U[0, 0.998],
    inclination ~ U[5, 85] deg, +-20 M field, h = 0.05 (r - r_+), <= 4096 steps; brick index i % 4."""
    rng = np.random.default_rng(int(seed))
    out = []
    for i in range(int(n)):
        npx = int(round(math.exp(rng.uniform(math.log(min_px), math.log(max_px)))))
        spin = float(rng.uniform(0.0, 0.998))
        inc = float(rng.uniform(5.0, 85.0))
        cam = Camera.spanning(r=1000.0, inclination_degrees=inc, half_width_m=20.0, width=npx, height=npx)
        out.append((Scene(spin=spin, camera=cam, width=npx, height=npx, step_mode=STEP_ADAPTIVE, step=0.05,
                          max_steps=4096, escape_radius=220.0), i % 4))
    return out
