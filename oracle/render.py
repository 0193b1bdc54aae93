"""FP64 oracle for the fused Kerr Stokes render (test infrastructure only).

One ray per pixel.  Each ray follows, in order:

* camera launch: kerr.py:188-233 (``camera_rays_vec``);
* geodesic steps of the scene's integrator, theta clamped after each: RK2
  midpoint, k1 = f(y), k2 = f(y + h/2 k1), y + h k2 (scheme of
  native/rocm/kerr_stokes_kernel.hip:13), or the reference tracer's classical
  RK4, y + (h/6)(k1 + 2 k2 + 2 k3 + k4) (kerr.py:306-313); RHS of
  kerr.py:285-303; h is the fixed affine step or h_frac * (r - r_+) taken
  from the state at the start of the step;
* after each step the transfer segment of accelerated_renderer.py:112-125:
  sample the brick at the coordinate midpoint of (previous, current) state
  (in FP64 like native_kernels.py's sampler, or with ``scene.numpy32_sampling``
  in float32 as the reference renderer's own sampler runs under NumPy),
  skip the segment when r or theta is outside the brick, otherwise
  ds = L * max(|delta(r, theta, phi)|, 1e-14), one RK2 Stokes step (or, with
  ``scene.exact_step``, the exact step of polarized_transfer.py:51-70), and
  tau += max(alpha_I, 0) ds;
* termination of kerr.py:357-377: a non-finite state ends the ray (the
  reference appends that state to the trace, so its transfer segment is still
  applied, with IEEE NaN propagation); capture when r <= r_+ (1 + 2e-4);
  escape when r > escape_radius and r grew; max_steps; and the optical-depth
  cap tau >= max_optical_depth (accelerated_renderer.py:123-125), after which
  the Stokes vector can no longer change;
* ``nan_to_num`` with NaN and +-inf mapped to 0 (accelerated_renderer.py:126).

The trace-then-transfer order of the reference is fused here (the transfer of
segment i only needs states i-1 and i); the Stokes result is the same.  The
sampler works in FP64 on the float32 grid and coefficient values, like
native_kernels.py's FP64 sampler.

Status codes (low 3 bits): 0 max_steps, 1 captured, 2 escaped, 3 invalid,
4 optically thick.  Bit 3 (value 8) marks a ray whose steps could not
follow the geodesic through a coordinate pole: the theta clamp of kerr.py:35-36
fired at a stage or end state, or one step changed phi by more than 1 rad
(d(phi)/dlambda = L / (Sigma sin^2 theta) blows up within a step of the axis;
regular steps move phi by well under 1 rad).  Such rays pick up huge ds and
their Stokes vectors are not meaningful in any precision, so error statistics
exclude them and report how many there were.
"""

from __future__ import annotations

import math
import os

import numpy as np

from . import kerr, micro

ST_MAX_STEPS, ST_CAPTURED, ST_ESCAPED, ST_INVALID, ST_OPAQUE = 0, 1, 2, 3, 4
ST_POLE_CLAMPED = 8
POLE_DPHI = 1.0   # rad per step


def _pymax(a, b):
    """Python's max(a, b) for arrays: b if b > a else a (keeps NaN in a)."""
    return np.where(b > a, b, a)


def _rk2_stokes(S, c, ds):
    """lib.rs:210-225 on (n,4) states with per-row ds."""
    d = ds[:, None]
    k1 = micro.stokes_rhs(S, c)
    k2 = micro.stokes_rhs(S + 0.5 * d * k1, c)
    return S + d * k2


def geodesic_step_vec(r, th, ph, pr, pth, pt, pph, h, a: float, rk4: bool):
    """One RK2 midpoint or RK4 step over arrays of rays.  Returns the end state (r, raw theta, phi, p_r,
    p_theta) and the theta of every intermediate stage (the states the RHS saw besides the start)."""
    k1 = kerr.geodesic_rhs_vec(r, th, pr, pth, pt, pph, a)
    hh = 0.5 * h
    th2 = th + hh * k1[1]
    k2 = kerr.geodesic_rhs_vec(r + hh * k1[0], th2, pr + hh * k1[3], pth + hh * k1[4], pt, pph, a)
    if not rk4:
        return (r + h * k2[0], th + h * k2[1], ph + h * k2[2], pr + h * k2[3], pth + h * k2[4]), (th2,)
    th3 = th + hh * k2[1]
    k3 = kerr.geodesic_rhs_vec(r + hh * k2[0], th3, pr + hh * k2[3], pth + hh * k2[4], pt, pph, a)
    th4 = th + h * k3[1]
    k4 = kerr.geodesic_rhs_vec(r + h * k3[0], th4, pr + h * k3[3], pth + h * k3[4], pt, pph, a)
    w = h / 6.0
    end = tuple(y + w * (k1[j] + 2.0 * k2[j] + 2.0 * k3[j] + k4[j]) for j, y in enumerate((r, th, ph, pr, pth)))
    return end, (th2, th3, th4)


def render_rays(scene, brick, pix: np.ndarray | None = None):
    """Render the pixels ``pix`` (flat indices y*W + x; all pixels when None).

    Returns dict(iquv=(n,4) float64, steps=(n,) int32, status=(n,) int8).
    """
    a = float(scene.spin)
    kerr.validate_spin(a)
    W, H = int(scene.width), int(scene.height)
    if pix is None:
        pix = np.arange(W * H, dtype=np.int64)
    pix = np.asarray(pix, dtype=np.int64)
    n = pix.size
    px, py = pix % W, pix // W
    ndc_x = 2.0 * (px + 0.5) / W - 1.0
    ndc_y = 1.0 - 2.0 * (py + 0.5) / H
    r, th, ph, pr, pth, pt, pph = kerr.camera_rays_vec(scene.camera, a, ndc_x, ndc_y, W / H)

    rg = brick.r.astype(np.float64)
    tg = brick.theta.astype(np.float64)
    pg = brick.phi.astype(np.float64)
    coeffs = brick.coeffs.astype(np.float64)
    r_plus = kerr.horizon_radius(a)
    r_cap = r_plus * (1.0 + 2.0e-4)
    adaptive = int(scene.step_mode) == 1
    h0 = float(scene.step)
    L = float(scene.affine_to_length_cm)
    tau_max = float(scene.max_optical_depth)
    r_esc = float(scene.escape_radius)
    exact = bool(getattr(scene, "exact_step", False))
    rk4 = getattr(scene, "integrator", "rk2") == "rk4"
    numpy32 = bool(getattr(scene, "numpy32_sampling", False))
    if numpy32:
        coeffs32 = brick.coeffs.astype(np.float32)

    S_out = np.zeros((n, 4))
    steps_out = np.zeros(n, dtype=np.int32)
    status_out = np.zeros(n, dtype=np.int8)

    idx = np.arange(n)
    S = np.zeros((n, 4))
    tau = np.zeros(n)
    lo, hi = kerr.EPS_THETA, math.pi - kerr.EPS_THETA
    clamped = np.zeros(n, dtype=bool)
    for i in range(1, int(scene.max_steps) + 1):
        if idx.size == 0:
            break
        h = h0 * (r - r_plus) if adaptive else np.full(r.shape, h0)
        (nr_, nth_raw, nph, npr, npth), stages = geodesic_step_vec(r, th, ph, pr, pth, pt, pph, h, a, rk4)
        nth = kerr.safe_theta(nth_raw)
        for t_ in stages + (nth_raw,):
            clamped |= (t_ < lo) | (t_ > hi)
        clamped |= np.abs(nph - ph) > POLE_DPHI
        finite = np.isfinite(nr_) & np.isfinite(nth) & np.isfinite(nph) & np.isfinite(npr) & np.isfinite(npth)

        # transfer segment (accelerated_renderer.py:113-125)
        with np.errstate(invalid="ignore", over="ignore"):
            mr, mth, mph = 0.5 * (r + nr_), 0.5 * (th + nth), 0.5 * (ph + nph)
            if numpy32:
                c, valid = micro.sample_vec_numpy32(coeffs32, brick.r, brick.theta, brick.phi, mr, mth, mph)
            else:
                c, valid = micro.sample_vec(coeffs, rg, tg, pg, mr, mth, mph)
            dr, dth, dph = nr_ - r, nth - th, nph - ph
            dlam = np.sqrt(dr * dr + dth * dth + dph * dph)
            ds = _pymax(L * _pymax(dlam, 1.0e-14), 0.0)
            if np.any(valid):
                if exact:
                    S[valid] = micro.stokes_exact_step(S[valid], c[valid], ds[valid])
                else:
                    S[valid] = _rk2_stokes(S[valid], c[valid], ds[valid])
                tau[valid] = tau[valid] + _pymax(c[valid, 4], 0.0) * ds[valid]

        status = np.full(idx.size, -1, dtype=np.int8)
        status[~finite] = ST_INVALID
        live = status < 0
        status[live & (nr_ <= r_cap)] = ST_CAPTURED
        live = status < 0
        status[live & (nr_ > r_esc) & (nr_ > r)] = ST_ESCAPED
        live = status < 0
        status[live & (tau >= tau_max)] = ST_OPAQUE
        done = status >= 0
        if i == int(scene.max_steps):
            status[~done] = ST_MAX_STEPS
            done[:] = True
        if np.any(done):
            d = idx[done]
            S_out[d] = S[done]
            steps_out[d] = i
            status_out[d] = status[done] | np.where(clamped[done], ST_POLE_CLAMPED, 0).astype(np.int8)
        keep = ~done
        idx = idx[keep]
        r, th, ph, pr, pth = nr_[keep], nth[keep], nph[keep], npr[keep], npth[keep]
        pt, pph, S, tau, clamped = pt[keep], pph[keep], S[keep], tau[keep], clamped[keep]

    S_out = np.nan_to_num(S_out, nan=0.0, posinf=0.0, neginf=0.0)
    return {"iquv": S_out, "steps": steps_out, "status": status_out}


def render_image(scene, brick):
    """Full image (H, W, 4) plus steps and status maps."""
    out = render_rays(scene, brick)
    H, W = scene.height, scene.width
    return {"iquv": out["iquv"].reshape(H, W, 4), "steps": out["steps"].reshape(H, W),
            "status": out["status"].reshape(H, W)}


def _band_worker(args):
    scene, brick, pix = args
    return render_rays(scene, brick, pix)


def render_rays_parallel(scene, brick, pix: np.ndarray, processes: int | None = None):
    """Split ``pix`` into contiguous bands and render them on a process pool (CPU baseline)."""
    import multiprocessing as mp

    procs = processes or len(os.sched_getaffinity(0))
    bands = [b for b in np.array_split(np.asarray(pix, dtype=np.int64), procs) if b.size]
    ctx = mp.get_context("fork")
    with ctx.Pool(len(bands)) as pool:
        parts = pool.map(_band_worker, [(scene, brick, b) for b in bands])
    return {k: np.concatenate([p[k] for p in parts]) for k in ("iquv", "steps", "status")}
