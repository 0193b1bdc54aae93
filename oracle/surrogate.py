"""FP64 oracle for the learned surrogate (test infrastructure only).

The reference has no learned part; these functions state what the HIP surrogate
kernels compute so the tests can check them:

* ``mlp_forward`` / ``mlp_backward``: x -> [Linear, SiLU] x depth -> Linear with
  the weighted squared-error loss L = sum_{i,o} w_o (y_io - t_io)^2 / rows, by
  the chain rule in FP64.  With ``bf16=True`` the values the GPU rounds to
  bfloat16 (inputs, weights, hidden activations, the back-propagated
  pre-activation gradients) are rounded the same way (round to nearest even),
  so only FP32 accumulation order separates the two.
* ``adam``: Kingma & Ba's update with bias correction (beta1 0.9, beta2 0.999,
  eps 1e-8).
* ``knots``: the FP64 RK2 geodesic of ``oracle.render`` (camera of kerr.py:188-233,
  RHS of kerr.py:285-303, termination of kerr.py:357-377) recorded from the start
  of the first step whose midpoint radius lies in the brick's radial range (and
  no longer once a pole clamp fired), with the ZAMO redshift
  g = E alpha / (E - omega L) per knot.
* ``knot_transfer``: the Stokes transfer of accelerated_renderer.py:113-125 over
  a knot sequence, ended by the rules the HIP kernel uses.

Parameters use the padded layout of the HIP MLP: a list of (W [fi_pad, fo_pad],
b [fo_pad]) per layer, n_in and n_out padded to multiples of 8 with zeros.
"""

from __future__ import annotations

import math

import numpy as np

from . import kerr, micro
from .render import _pymax, _rk2_stokes

HEADS = 3   # captured, flag, recorded states


# ---- MLP ----------------------------------------------------------------------

def bf16_round(x):
    """float -> nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def pad8(n: int) -> int:
    return (int(n) + 7) // 8 * 8


def layer_shapes(n_in: int, width: int, depth: int, n_out: int):
    dims = [pad8(n_in)] + [int(width)] * int(depth) + [pad8(n_out)]
    return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]


def flatten(params) -> np.ndarray:
    """List of (W, b) -> the flat padded parameter vector of the HIP MLP."""
    return np.concatenate([np.concatenate([np.asarray(W, np.float64).ravel(), np.asarray(b, np.float64).ravel()])
                           for W, b in params])


def unflatten(vec, shapes):
    out, off = [], 0
    for fi, fo in shapes:
        W = np.asarray(vec[off:off + fi * fo], dtype=np.float64).reshape(fi, fo)
        off += fi * fo
        b = np.asarray(vec[off:off + fo], dtype=np.float64)
        off += fo
        out.append((W, b))
    return out


def _pad_x(x, fi):
    x = np.asarray(x, dtype=np.float64)
    if x.shape[1] < fi:
        x = np.concatenate([x, np.zeros((x.shape[0], fi - x.shape[1]))], axis=1)
    return x


def mlp_forward(params, x, bf16: bool = False):
    """Returns (y [rows, fo_pad], cache) with cache = (inputs per layer, pre-activations per hidden layer)."""
    rnd = bf16_round if bf16 else (lambda v: v)
    h = rnd(_pad_x(x, params[0][0].shape[0]))
    inputs, zs = [], []
    for l, (W, b) in enumerate(params):
        inputs.append(h)
        z = h @ rnd(W) + (_f32(b) if bf16 else b)
        if l < len(params) - 1:
            z = _f32(z) if bf16 else z
            zs.append(z)
            h = rnd(silu(z))
        else:
            h = _f32(z) if bf16 else z
    return h, (inputs, zs)


def mlp_loss_grad(params, x, t, w, bf16: bool = False, row_w=None, masked: int = 0):
    """Loss and its gradient (list of (dW, db)) for targets t [rows, n_out] and weights w [n_out];
    row_w [rows] (optional) multiplies the weights of the first ``masked`` outputs of each row."""
    y, (inputs, zs) = mlp_forward(params, x, bf16)
    rows, n_out = np.asarray(t).shape
    d = y[:, :n_out] - np.asarray(t, dtype=np.float64)
    w = np.broadcast_to(np.asarray(w, dtype=np.float64), (rows, n_out)).copy()
    if row_w is not None:
        w[:, :masked] *= np.asarray(row_w, dtype=np.float64)[:, None]
    loss = float(np.sum(w * d * d) / rows)
    dy = np.zeros_like(y)
    dy[:, :n_out] = 2.0 * w * d / rows
    rnd = bf16_round if bf16 else (lambda v: v)
    grads = [None] * len(params)
    dz_f = _f32(dy) if bf16 else dy   # FP32 gradient (bias), BF16 copy feeds the GEMMs
    dz_b = rnd(dz_f)
    for l in range(len(params) - 1, -1, -1):
        W, _ = params[l]
        grads[l] = (inputs[l].T @ dz_b, dz_f.sum(axis=0))
        if l == 0:
            break
        g = (dz_b @ rnd(W).T) * dsilu(zs[l - 1])
        dz_f = _f32(g) if bf16 else g
        dz_b = rnd(dz_f)
    return loss, grads


def adam(params, grads, m, v, t: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) on flat vectors; returns (params, m, v)."""
    m = b1 * m + (1.0 - b1) * grads
    v = b2 * v + (1.0 - b2) * grads * grads
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


# ---- knots --------------------------------------------------------------------

def zamo_redshift(a, r, th, E, L):
    """g = E alpha / (E - omega L) for a zero-angular-momentum observer at (r, theta)."""
    s, c = np.sin(th), np.cos(th)
    sig = r * r + a * a * c * c
    dlt = r * r - 2.0 * r + a * a
    A = (r * r + a * a) ** 2 - a * a * dlt * s * s
    lapse = np.sqrt(np.maximum(sig * dlt / A, 0.0))
    omega = 2.0 * a * r / A
    return E * lapse / (E - omega * L)


def _camera_state(cam_r, alpha, beta, a, cos_i):
    """Initial state of the ray with ZAMO-frame direction (-1, -beta/cam_r, alpha/cam_r) (kerr.py:188-233)."""
    th = float(kerr.safe_theta(math.acos(min(max(float(cos_i), -1.0), 1.0))))
    e = kerr.zamo_tetrad(cam_r, th, a)
    g = kerr.metric_covariant(cam_r, th, a)
    n = np.array([-1.0, -float(beta) / cam_r, float(alpha) / cam_r])
    n /= np.linalg.norm(n)
    p_con = e[0] + n[0] * e[1] + n[1] * e[2] + n[2] * e[3]
    p = g @ p_con
    return th, float(p[1]), float(p[2]), float(p[0]), float(p[3])


def knots(params, cam_r, step, adaptive, max_steps, r_esc, r_lo, r_hi, K):
    """Knot rows (n, 4K + 3) and total step counts for rays params (n, 4) = (alpha, beta, a, cos i)."""
    params = np.asarray(params, dtype=np.float64)
    out = np.zeros((params.shape[0], 4 * K + HEADS))
    steps_out = np.zeros(params.shape[0], dtype=np.int32)
    lo, hi = kerr.EPS_THETA, math.pi - kerr.EPS_THETA
    for i, (alpha, beta, a, ci) in enumerate(params):
        th, pr, pth, pt, pph = _camera_state(cam_r, alpha, beta, a, ci)
        E, L = -pt, pph
        r_plus = kerr.horizon_radius(a)
        r_cap = r_plus * (1.0 + 2.0e-4)
        r, ph = float(cam_r), 0.0
        rec = []
        entered = overflow = clamped = False
        status, n = 0, 0
        while True:
            h = step * (r - r_plus) if adaptive else step
            k1 = kerr.geodesic_rhs_vec(r, th, pr, pth, pt, pph, a)
            mth = th + 0.5 * h * k1[1]
            k2 = kerr.geodesic_rhs_vec(r + 0.5 * h * k1[0], mth, pr + 0.5 * h * k1[3], pth + 0.5 * h * k1[4], pt, pph, a)
            nr = r + h * k2[0]
            nt_raw = th + h * k2[1]
            nt = float(kerr.safe_theta(nt_raw))
            nph = ph + h * k2[2]
            npr, npth = pr + h * k2[3], pth + h * k2[4]
            clamped = clamped or mth < lo or mth > hi or nt_raw < lo or nt_raw > hi or abs(nph - ph) > 1.0
            fin = all(math.isfinite(v) for v in (nr, nt, nph, npr, npth))
            rm = 0.5 * (r + nr)
            inb = r_lo <= rm <= r_hi
            if inb and not entered:
                entered = True
                rec.append((r, th, ph))
            if entered and fin and not clamped:
                if len(rec) < K:
                    rec.append((nr, nt, nph))
                elif inb:
                    overflow = True
            r_prev = r
            r, th, ph, pr, pth = nr, nt, nph, npr, npth
            n += 1
            if not fin:
                status = 3
                break
            if nr <= r_cap:
                status = 1
                break
            if nr > r_esc and nr > r_prev:
                status = 2
                break
            if n >= max_steps:
                status = 0
                break
        n_rec = len(rec)
        if not rec:
            rec.append((r, th, ph))
        while len(rec) < K:
            rec.append(rec[-1])
        row = out[i]
        for j, (kr, kt, kp) in enumerate(rec):
            row[4 * j:4 * j + 4] = (kr, kt, kp, zamo_redshift(a, kr, kt, E, L))
        row[4 * K] = 1.0 if status == 1 else 0.0
        row[4 * K + 1] = 1.0 if (overflow or clamped or status in (0, 3)) else 0.0
        row[4 * K + 2] = n_rec
        steps_out[i] = n
    return out, steps_out


def knot_transfer(rows, K, brick, L, tau_max, r_cap, cam_phi=0.0):
    """Stokes IQUV (n, 4), status (n,) and segment counts (n,) of knot rows (n, 4K + 3).

    Segment j joins knots j and j+1; it is skipped unless its midpoint samples the
    brick, as in accelerated_renderer.py:113-125.  A ray ends when a segment's
    midpoint lies outside the brick's radial range while r moves away from it
    (captured when moving in, escaped when moving out), when a knot reaches
    r <= r_cap, or when tau >= tau_max; otherwise after K - 1 segments, with the
    captured head deciding the status.
    """
    rows = np.asarray(rows, dtype=np.float64)
    rg, tg, pg = (brick.r.astype(np.float64), brick.theta.astype(np.float64), brick.phi.astype(np.float64))
    coeffs = brick.coeffs.astype(np.float64)
    r_lo, r_hi = rg[0], rg[-1]
    n = rows.shape[0]
    S_out = np.zeros((n, 4))
    status = np.zeros(n, dtype=np.int32)
    segs = np.zeros(n, dtype=np.int32)
    for i in range(n):
        kn = rows[i, :4 * K].reshape(K, 4)
        r = kn[:, 0]
        th = kerr.safe_theta(kn[:, 1])
        ph = kn[:, 2] + cam_phi
        S = np.zeros((1, 4))
        tau = 0.0
        code = -1
        seg = 0
        for j in range(1, K):
            rm = 0.5 * (r[j - 1] + r[j])
            if (rm > r_hi and r[j] > r[j - 1]) or (rm < r_lo and r[j] < r[j - 1]):
                code = 1 if r[j] < r[j - 1] else 2
                break
            c, valid = micro.sample_vec(coeffs, rg, tg, pg, np.array([rm]), np.array([0.5 * (th[j - 1] + th[j])]),
                                        np.array([0.5 * (ph[j - 1] + ph[j])]))
            if valid[0]:
                d = np.array([r[j] - r[j - 1], th[j] - th[j - 1], ph[j] - ph[j - 1]])
                ds = _pymax(L * _pymax(math.sqrt(float(d @ d)), 1.0e-14), 0.0)
                S = _rk2_stokes(S, c, np.array([ds]))
                tau += max(float(c[0, 4]), 0.0) * ds
            seg += 1
            if r[j] <= r_cap:
                code = 1
                break
            if tau >= tau_max:
                code = 4
                break
        if code < 0:
            code = 1 if rows[i, 4 * K] >= 0.5 else 2
        S_out[i] = np.nan_to_num(S[0], nan=0.0, posinf=0.0, neginf=0.0)
        status[i] = code
        segs[i] = seg
    return S_out, status, segs
