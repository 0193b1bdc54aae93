"""Pins the surrogate oracle (oracle/surrogate.py) on the CPU.

* bf16 rounding: against torch's float32 -> bfloat16 conversion and hand-worked ties.
* MLP backward: against central finite differences of the FP64 loss.
* MLP forward: a one-path network in closed form.
* Adam: the first step moves every weight by lr * g / (|g| + eps) (m_hat = g, v_hat = g^2).
* knots: the Schwarzschild redshift sqrt(1 - 2/r) of a static observer, and the transfer over
  the recorded knots reproduces the oracle's own RK2 render for every unflagged ray.
"""

import math

import numpy as np
import pytest

from bhgrrt import scenes
from oracle import kerr, surrogate as osur
from oracle import render as orender


def test_bf16_round_matches_torch_and_ties_to_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=4096) * 10.0 ** rng.integers(-6, 6, 4096), [0.0, -0.0, 1.0, -3.5]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(osur.bf16_round(x), ref)
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7 (odd mantissa): ties go to 1
    assert osur.bf16_round(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3 * 2^-8 lies halfway between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    assert osur.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def _tiny(seed=0, n_in=4, width=8, depth=2, n_out=5):
    rng = np.random.default_rng(seed)
    shapes = osur.layer_shapes(n_in, width, depth, n_out)
    params = []
    for l, (fi, fo) in enumerate(shapes):
        W = np.zeros((fi, fo))
        ri = n_in if l == 0 else fi
        ro = n_out if l == len(shapes) - 1 else fo
        W[:ri, :ro] = rng.normal(size=(ri, ro)) * 0.7
        b = np.zeros(fo)
        b[:ro] = rng.normal(size=ro) * 0.1
        params.append((W, b))
    return params, shapes


def test_mlp_backward_matches_finite_differences():
    params, shapes = _tiny()
    rng = np.random.default_rng(3)
    x = rng.normal(size=(16, 4))
    t = rng.normal(size=(16, 5))
    w = rng.uniform(0.5, 2.0, size=5)
    loss, grads = osur.mlp_loss_grad(params, x, t, w)
    flat, gflat = osur.flatten(params), osur.flatten(grads)
    eps = 1e-6
    idx = rng.choice(flat.size, 60, replace=False)
    for i in idx:
        p1, p2 = flat.copy(), flat.copy()
        p1[i] += eps
        p2[i] -= eps
        l1, _ = osur.mlp_loss_grad(osur.unflatten(p1, shapes), x, t, w)
        l2, _ = osur.mlp_loss_grad(osur.unflatten(p2, shapes), x, t, w)
        fd = (l1 - l2) / (2 * eps)
        assert abs(fd - gflat[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, gflat[i])
    # padded rows of the first layer and padded output columns get zero gradient
    assert np.all(grads[0][0][4:] == 0.0)
    assert np.all(grads[-1][0][:, 5:] == 0.0) and np.all(grads[-1][1][5:] == 0.0)


def test_mlp_forward_one_path_closed_form():
    params, _ = _tiny()
    params = [(np.zeros_like(W), np.zeros_like(b)) for W, b in params]
    params[0][0][1, 2] = 0.5       # x1 -> hidden unit 2
    params[0][1][2] = -0.25
    params[1][0][2, 3] = 2.0       # -> unit 3 of layer 2
    params[2][0][3, 4] = -1.5      # -> output 4
    params[2][1][4] = 0.125
    x = np.array([[0.0, 1.2, 0.0, 0.0], [0.0, -3.0, 0.0, 0.0]])
    y, _ = osur.mlp_forward(params, x)
    s = lambda z: z / (1.0 + math.exp(-z))
    for i in range(2):
        h1 = s(0.5 * x[i, 1] - 0.25)
        h2 = s(2.0 * h1)
        assert y[i, 4] == pytest.approx(-1.5 * h2 + 0.125, rel=1e-14)
        assert np.all(y[i, :4] == 0.0)


def test_adam_first_step_is_lr_sign():
    rng = np.random.default_rng(5)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p1, m, v = osur.adam(p, g, np.zeros(100), np.zeros(100), 1, 1e-3)
    assert np.allclose(p - p1, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    # second step with the same gradient keeps the same magnitude (bias-corrected averages of a constant)
    p2, _, _ = osur.adam(p1, g, m, v, 2, 1e-3)
    assert np.allclose(p1 - p2, 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=0)


def test_schwarzschild_knot_redshift_is_static_observer():
    params = np.array([[3.0, -4.0, 0.0, math.cos(math.radians(40.0))], [12.0, 2.0, 0.0, 0.5]])
    rows, _ = osur.knots(params, 1000.0, 0.05, True, 4096, 220.0, 2.1, 40.0, 32)
    for row in rows:
        kn = row[:4 * 32].reshape(32, 4)
        n = int(row[4 * 32 + 2])
        assert n > 2
        r, g = kn[:n, 0], kn[:n, 3]
        assert np.allclose(g, np.sqrt(1.0 - 2.0 / r), rtol=1e-12)
    # far from the hole the ZAMO redshift of any spin tends to 1
    g_far = osur.zamo_redshift(0.9, 1.0e6, 1.0, 1.0, 3.0)
    assert abs(g_far - 1.0) < 1e-5


def test_knot_transfer_reproduces_the_rk2_render():
    """Knots recorded from the oracle's own RK2 geodesics, transferred by knot_transfer, give the oracle
    render's Stokes vector and status for every unflagged ray whose brick part fits in K knots."""
    sc = scenes.scale_plain_scene(32, 32)
    br = scenes.scale_brick(32, 16, 32)
    K = 128
    cam = sc.camera
    tan = math.tan(0.5 * math.radians(cam.fov_y_degrees))
    pix = np.arange(0, 32 * 32, 7)
    px, py = pix % 32, pix // 32
    ndc_x, ndc_y = 2.0 * (px + 0.5) / 32 - 1.0, 1.0 - 2.0 * (py + 0.5) / 32
    params = np.stack([cam.r * ndc_x * tan, cam.r * ndc_y * tan, np.full(pix.size, sc.spin),
                       np.full(pix.size, math.cos(cam.theta))], axis=1)
    rows, steps = osur.knots(params, cam.r, sc.step, True, sc.max_steps, sc.escape_radius,
                             float(np.float32(br.r[0])), float(np.float32(br.r[-1])), K)
    r_cap = kerr.horizon_radius(sc.spin) * (1.0 + 2.0e-4)
    S, status, segs = osur.knot_transfer(rows, K, br, sc.affine_to_length_cm, sc.max_optical_depth, r_cap)
    ref = orender.render_rays(sc, br, pix)
    ok = (rows[:, 4 * K + 1] == 0) & ((ref["status"] & orender.ST_POLE_CLAMPED) == 0)
    assert ok.sum() > 0.8 * pix.size
    I, I0 = S[ok, 0], ref["iquv"][ok, 0]
    assert np.abs(I - I0).sum() <= 1e-9 * np.abs(I0).sum()
    assert np.allclose(S[ok], ref["iquv"][ok], rtol=1e-8, atol=1e-12 * np.abs(I0).max())
    # the transfer ends rays the way the RK2 render does (captured / escaped / opaque)
    assert np.array_equal(status[ok], ref["status"][ok] & 7)
    # total RK2 steps of the knot tracer match the render for rays that were not cut by opacity
    thin = ok & (ref["status"] != orender.ST_OPAQUE)
    assert np.array_equal(steps[thin], ref["steps"][thin])
