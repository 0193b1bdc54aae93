"""Write tests/golden/accelerated_vectors.npz from the reference's own CPU renderer.

Run in a container that holds the reference (never on the GPU box; nothing imports this file):

    PYTHONPATH=/root/reference/python python -B tests/golden/make_accelerated_vectors.py

For every pixel of four small tiles it calls the reference's
render_tile_stokes_bricks (python/blackhole_sim/accelerated_renderer.py:129-150:
camera launch, trace_kerr_null_geodesic with rk4_step, then
integrate_stokes_from_bricks_trace) on a one-pixel tile, and
trace_kerr_null_geodesic (kerr.py:337-384) once more with the same arguments
for the trace status and its number of steps (states - 1).  The cases:

* default: AcceleratedRenderConfig() (accelerated_renderer.py:37-47): 160x90,
  camera r=55, i=65 deg, fov 32 deg, step 0.06, 5000 steps, escape 220, RK2
  Stokes step; two 16x12 tiles (image centre, and off-centre across the disk);
* exact: the same config with use_exact_matrix_step=True (polarized_transfer.py:51-70),
  an 8x6 tile at the centre;
* cli: the defaults of accelerated_cli.py:26-33 (64x36, a=0.85, camera r=55,
  i=65 deg with LocalCamera's default fov 35 deg, step 0.08, 700 steps), a 16x12 tile.

The spin is 0.85 (accelerated_cli.py:28).  The brick is a seeded 12x8x12
synthetic brick (r geometric over [1.05 r_+, 40], theta over [0, pi], phi over
[0, 2 pi) without endpoint; lognormal emissivity and absorption with a thin
disk, polarised fractions, Faraday terms), stored in the fixture as float32
with its grids, so the fixture is self-contained.
"""

from __future__ import annotations

import math
import multiprocessing as mp
from dataclasses import replace
from pathlib import Path
from types import SimpleNamespace

import numpy as np

from blackhole_sim.accelerated_renderer import AcceleratedRenderConfig, render_tile_stokes_bricks
from blackhole_sim.coefficient_bricks import CoefficientBrickGrid
from blackhole_sim.kerr import LocalCamera, camera_ray_initial_state, horizon_radius, trace_kerr_null_geodesic
from blackhole_sim.polarized_transfer import PolarizedTransferConfig
from blackhole_sim.tiles import RenderTile

OUT = Path(__file__).resolve().parent / "accelerated_vectors.npz"
SPIN = 0.85
STATUS = {"max_steps": 0, "captured": 1, "escaped": 2, "invalid": 3}


def make_brick(nr=12, nt=8, nph=12, seed=20241204):
    rng = np.random.default_rng(seed)
    rp = horizon_radius(SPIN)
    r = np.geomspace(1.05 * rp, 40.0, nr)
    th = np.linspace(0.0, math.pi, nt)
    ph = np.linspace(0.0, 2.0 * math.pi, nph, endpoint=False)
    shape = (nr, nt, nph)
    R = r[:, None, None] * np.ones(shape)
    TH = th[None, :, None] * np.ones(shape)
    j_i = R ** -2 * rng.lognormal(0.0, 0.5, shape) * (0.05 + np.exp(-np.cos(TH) ** 2 / 0.1))
    a_i = R ** -2 * rng.lognormal(0.0, 0.5, shape)
    pl = rng.uniform(0.0, 0.7, shape)
    chi = rng.uniform(0.0, math.pi, shape)
    pc = rng.uniform(-0.7, 0.7, shape)
    c = np.stack([j_i, pl * j_i * np.cos(2 * chi), pl * j_i * np.sin(2 * chi), pc * j_i,
                  a_i, pl * a_i * np.cos(2 * chi), pl * a_i * np.sin(2 * chi), pc * a_i,
                  rng.normal(0, 0.1, shape) * a_i, rng.normal(0, 0.1, shape) * a_i, rng.normal(0, 0.1, shape) * a_i],
                 axis=-1)
    return r.astype(np.float32), th.astype(np.float32), ph.astype(np.float32), c.astype(np.float32)


def cases():
    base = AcceleratedRenderConfig()
    exact = replace(base, transfer=PolarizedTransferConfig(use_exact_matrix_step=True))
    cli = AcceleratedRenderConfig(width=64, height=36, camera=LocalCamera.from_degrees(r=55.0, inclination_degrees=65.0),
                                  step=0.08, max_steps=700)
    return [("default_a", base, (72, 39, 88, 51)), ("default_b", base, (104, 20, 120, 32)),
            ("exact", exact, (76, 42, 84, 48)), ("cli", cli, (24, 12, 40, 24))]


BRICK = None


def pixel(args):
    name, cfg, i, j = args
    snap = SimpleNamespace(spin_a=SPIN)   # render_tile_stokes_bricks reads only snapshot.spin_a
    stokes = render_tile_stokes_bricks(cfg, snap, BRICK, RenderTile(0, y0=j, x0=i, x1=i + 1, y1=j + 1))[0, 0]
    ndc_x = 2.0 * (i + 0.5) / cfg.width - 1.0
    ndc_y = 1.0 - 2.0 * (j + 0.5) / cfg.height
    y0, p_t, p_phi = camera_ray_initial_state(cfg.camera, SPIN, ndc_x, ndc_y, cfg.width / cfg.height)
    tr = trace_kerr_null_geodesic(y0, p_t, p_phi, SPIN, step=cfg.step, max_steps=cfg.max_steps,
                                  escape_radius=cfg.escape_radius, store_stride=1)
    return stokes, STATUS[tr.status], tr.states.shape[0] - 1


def main():
    global BRICK
    r, th, ph, c = make_brick()
    BRICK = CoefficientBrickGrid(r=r, theta=th, phi=ph, coeffs=c, nu_hz=230e9, spin_a=SPIN)
    out = {"spin": np.array(SPIN), "brick_r": r, "brick_theta": th, "brick_phi": ph, "brick_coeffs": c}
    with mp.get_context("fork").Pool() as pool:
        for name, cfg, (x0, y0, x1, y1) in cases():
            jobs = [(name, cfg, i, j) for j in range(y0, y1) for i in range(x0, x1)]
            res = pool.map(pixel, jobs)
            h, w = y1 - y0, x1 - x0
            out[f"{name}_stokes"] = np.array([s for s, _, _ in res]).reshape(h, w, 4)
            out[f"{name}_status"] = np.array([s for _, s, _ in res], dtype=np.int8).reshape(h, w)
            out[f"{name}_steps"] = np.array([n for _, _, n in res], dtype=np.int32).reshape(h, w)
            out[f"{name}_tile"] = np.array([x0, y0, x1, y1])
            cam = cfg.camera
            out[f"{name}_config"] = np.array([cfg.width, cfg.height, cam.r, cam.theta, cam.phi, cam.fov_y_degrees,
                                              cam.roll_degrees, cfg.step, cfg.max_steps, cfg.escape_radius,
                                              float(cfg.transfer.use_exact_matrix_step),
                                              cfg.transfer.max_optical_depth,
                                              cfg.transfer.affine_to_length_cm or 1.0])
            print(name, "statuses", np.bincount(out[f"{name}_status"].ravel(), minlength=4),
                  "steps", out[f"{name}_steps"].min(), out[f"{name}_steps"].max(), flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
