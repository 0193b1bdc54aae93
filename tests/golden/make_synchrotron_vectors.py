"""Write tests/golden/synchrotron_vectors.npz from the reference's own coefficient precompute.

Run in a container that holds the reference (never on the GPU box; nothing imports this file):

    PYTHONPATH=/root/reference/python python -B tests/golden/make_synchrotron_vectors.py

It calls python/blackhole_sim/coefficient_bricks.py:118-152 (precompute_coefficient_bricks) on the
analytic torus fixtures of the reference's tests (tests/test_coefficient_bricks.py:16-36) and of
its benchmark (benchmark.py:43-68), HybridSynchrotronCoefficients.coefficients
(synchrotron.py:315-325) on hand-picked plasma frames, and stores inputs and outputs as plain
arrays: the torus fields, the scalings, the kernel table (synchrotron.py:150-155) and the bricks.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

from blackhole_sim.calibration import PhysicalScaling
from blackhole_sim.coefficient_bricks import precompute_coefficient_bricks
from blackhole_sim.grmhd import generate_analytic_grmhd_torus
from blackhole_sim.physics import BlackHoleSystem
from blackhole_sim.synchrotron import HybridSynchrotronCoefficients, LocalPlasmaFrame, _synchrotron_kernel_table

OUT = Path(__file__).resolve().parent / "synchrotron_vectors.npz"


def torus_case(tag, snap, scaling, dtypes, stride, out):
    out[f"{tag}_r"], out[f"{tag}_theta"], out[f"{tag}_phi"] = snap.r, snap.theta, snap.phi
    out[f"{tag}_rho"], out[f"{tag}_theta_e"] = snap.rho, snap.theta_e
    out[f"{tag}_pressure"] = snap.pressure
    out[f"{tag}_b_con"] = snap.b_con
    out[f"{tag}_scaling"] = np.array([scaling.rho_cgs_per_code, scaling.b_gauss_per_code,
                                      scaling.mean_molecular_weight_per_electron])
    out[f"{tag}_stride"] = np.array(stride)
    for dt in dtypes:
        b = precompute_coefficient_bricks(snap, scaling, nu_hz=230e9, dtype=dt, stride=stride)
        out[f"{tag}_coeffs_{dt}"] = b.coeffs
        out[f"{tag}_grids_{dt}"] = np.concatenate([b.r, b.theta, b.phi])


def main() -> int:
    out: dict[str, np.ndarray] = {}
    torus_case("a", generate_analytic_grmhd_torus(spin_a=0.5, nr=5, ntheta=4, nphi=3),
               PhysicalScaling.from_mdot(BlackHoleSystem.m87_star(0.5), 1e-4), ("float64", "float32"), 1, out)
    torus_case("b", generate_analytic_grmhd_torus(spin_a=0.3, nr=6, ntheta=5, nphi=4),
               PhysicalScaling.from_mdot(BlackHoleSystem.sgr_a_star(0.3), 1e-8), ("float64", "float16"), 2, out)
    torus_case("c", generate_analytic_grmhd_torus(nr=8, ntheta=6, nphi=8),
               PhysicalScaling(mass_bh_g=1.0, distance_cm=1.0, rho_cgs_per_code=1.0e-18, b_gauss_per_code=30.0),
               ("float64", "float32"), 1, out)

    # plasma frames: n_e [cm^-3], theta_e, B [G], evpa [rad]; nu [Hz]
    rng = np.random.default_rng(20240611)
    frames = [(1e6, 10.0, 30.0, 0.3), (1e5, 1.0, 5.0, -1.2), (1e7, 50.0, 100.0, 2.5), (3e4, 0.05, 0.5, 0.0),
              (1e8, 0.2, 1000.0, 1.0), (1e3, 200.0, 0.1, -0.4), (1e6, 2e-3, 20.0, 0.7), (1e6, 1e-2, 20.0, 0.7),
              (0.0, 5.0, 10.0, 0.1), (1e6, 5.0, 0.0, 0.1)]
    for _ in range(22):
        frames.append((10 ** rng.uniform(2, 8), 10 ** rng.uniform(-1.5, 2), 10 ** rng.uniform(-1, 3), rng.uniform(-3, 3)))
    frames = np.array(frames)
    model = HybridSynchrotronCoefficients()
    rows = []
    for nu in (230e9, 86e9):
        for ne, th, b, evpa in frames:
            c = model.coefficients(LocalPlasmaFrame(ne, th, b, 0.0, evpa), nu)
            rows.append([c.j_i, c.j_q, c.j_u, c.j_v, c.alpha_i, c.alpha_q, c.alpha_u, c.alpha_v,
                         c.rho_v, c.rho_q, c.rho_u])
    out["frames"] = frames
    out["frame_nu"] = np.array([230e9, 86e9])
    out["frame_coeffs"] = np.array(rows).reshape(2, len(frames), 11)
    xs, F, G = _synchrotron_kernel_table()
    out["table_x"], out["table_F"], out["table_G"] = xs, F, G
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
