"""GPU parity of the surrogate kernels (BF16 MFMA MLP, Adam, geodesic knots, transfer over knots)."""


import numpy as np
import pytest

from bhgrrt import scenes
from bhgrrt import surrogate as bsur
from bhgrrt.render import PREC_FP32
from oracle import surrogate as osur

pytestmark = pytest.mark.gpu


def _params_for(mlp, seed, scale=1.0):
    p = mlp.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    out = []
    for l, (W, b) in enumerate(p):
        b = b.copy()
        ro = mlp.n_out if l == len(p) - 1 else b.size
        b[:ro] = rng.normal(size=ro).astype(np.float32) * 0.1
        out.append((W * np.float32(scale), b))
    return out


def _upload(renderer, a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    d = renderer.alloc(a.nbytes)
    d.from_host(a)
    return d


@pytest.mark.parametrize("rows,n_out", [(96, 20), (2048, 515)])
def test_mlp_forward_matches_bf16_oracle(renderer, rows, n_out):
    print(f"mlp forward rows={rows} n_out={n_out}", flush=True)
    mlp = bsur.MLP(renderer, 4, 256, 4, n_out)
    params = _params_for(mlp, 7)
    mlp.set_params(params)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, size=(rows, 4)).astype(np.float32)
    dx, dy = _upload(renderer, x), renderer.alloc(rows * n_out * 4)
    mlp.forward(dx, rows, dy)
    y = dy.to_host(np.empty((rows, n_out), np.float32)).astype(np.float64)
    ref, _ = osur.mlp_forward([(W.astype(np.float64), b.astype(np.float64)) for W, b in params], x, bf16=True)
    ref = ref[:, :n_out]
    # BF16 hidden activations: an FP32 difference can flip one activation's rounding (2^-8 relative),
    # so the bound is on the mean error and loose on the worst value
    assert np.abs(y - ref).sum() / np.abs(ref).sum() < 1e-3
    assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-2
    # transposed output ([n_out][rows]) holds the same values
    dyt = renderer.alloc(rows * n_out * 4)
    mlp.forward(dx, rows, dyt, transposed=True)
    yt = dyt.to_host(np.empty((n_out, rows), np.float32))
    assert np.array_equal(yt.T, y.astype(np.float32))
    # the FP64 network without bf16 rounding agrees to bf16 precision only
    ref64, _ = osur.mlp_forward([(W.astype(np.float64), b.astype(np.float64)) for W, b in params], x)
    assert np.abs(y - ref64[:, :n_out]).max() / np.abs(ref64).max() < 5e-2
    for d in (dx, dy, dyt):
        d.free()
    mlp.close()


@pytest.mark.parametrize("rows", [64, 8192])
def test_mlp_gradient_matches_oracle(renderer, rows):
    print(f"mlp gradient rows={rows}", flush=True)
    n_out = 515
    mlp = bsur.MLP(renderer, 4, 256, 4, n_out)
    params = _params_for(mlp, 3)
    mlp.set_params(params)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(rows, 4)).astype(np.float32)
    t = rng.normal(size=(rows, n_out)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=n_out).astype(np.float32)
    rw = (rng.uniform(size=rows) > 0.3).astype(np.float32)
    dx, dt, dw, drw = _upload(renderer, x), _upload(renderer, t), _upload(renderer, w), _upload(renderer, rw)
    loss, grads = mlp.gradient(dx, dt, dw, rows, drw, 300)
    p64 = [(W.astype(np.float64), b.astype(np.float64)) for W, b in params]
    loss_ref, g_ref = osur.mlp_loss_grad(p64, x, t, w, bf16=True, row_w=rw, masked=300)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(grads, g_ref)):
        eW = np.linalg.norm(gW - rW) / np.linalg.norm(rW)
        eb = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
        assert eW < 2e-3 and eb < 2e-3, (l, eW, eb)
    # against the FP64 chain rule without rounding: bf16-level agreement, same direction
    _, g64 = osur.mlp_loss_grad(p64, x, t, w, bf16=False, row_w=rw, masked=300)
    a, b = osur.flatten(grads), osur.flatten(g64)
    assert float(a @ b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.999
    # padded entries receive zero gradient
    assert np.all(grads[0][0][4:] == 0.0) and np.all(grads[-1][0][:, n_out:] == 0.0)
    for d in (dx, dt, dw, drw):
        d.free()
    mlp.close()


def test_adam_steps_match_oracle(renderer):
    print("adam two steps", flush=True)
    rows, n_out = 512, 64
    mlp = bsur.MLP(renderer, 4, 256, 4, n_out)
    params = _params_for(mlp, 9)
    mlp.set_params(params)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(rows, 4)).astype(np.float32)
    t = rng.normal(size=(rows, n_out)).astype(np.float32)
    w = np.ones(n_out, np.float32)
    dx, dt, dw = _upload(renderer, x), _upload(renderer, t), _upload(renderer, w)
    p = mlp.flat(params).astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in (1, 2):
        _, g = mlp.gradient(dx, dt, dw, rows)   # gradient at the current parameters (no update)
        mlp.train_step(dx, dt, dw, rows, 1e-3)
        p, m, v = osur.adam(p, mlp.flat(g).astype(np.float64), m, v, step, 1e-3)
        got = mlp.flat(mlp.get_params()).astype(np.float64)
        assert np.abs(got - p).max() < 2e-6, (step, np.abs(got - p).max())
    for d in (dx, dt, dw):
        d.free()
    mlp.close()


def test_training_reduces_the_loss(renderer):
    print("short training run", flush=True)
    rows, n_out = 4096, 32
    mlp = bsur.MLP(renderer, 4, 256, 4, n_out)
    mlp.set_params(mlp.init_params(1))
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(rows, 4)).astype(np.float32)
    # smooth targets: products of low-frequency functions of the inputs
    t = np.stack([np.sin(0.5 * (k % 4 + 1) * x[:, k % 4]) * np.cos(x[:, (k + 1) % 4]) + 0.1 * k / n_out
                  for k in range(n_out)], 1).astype(np.float32)
    dx, dt, dw = _upload(renderer, x), _upload(renderer, t), _upload(renderer, np.ones(n_out, np.float32))
    first = mlp.train_step(dx, dt, dw, rows, 1e-3, want_loss=True)
    for _ in range(300):
        last = mlp.train_step(dx, dt, dw, rows, 1e-3, want_loss=True)
    assert last < 0.05 * first, (first, last)
    for d in (dx, dt, dw):
        d.free()
    mlp.close()


def test_geodesic_knots_match_oracle(renderer):
    print("geodesic knots vs oracle", flush=True)
    cfg = bsur.SurrogateConfig(n_knots=64)
    br = scenes.scale_brick(32, 16, 32)
    params = bsur.sample_params(48, 77, cfg)
    dp = _upload(renderer, params)
    n_out = cfg.n_out
    drows, dsteps = renderer.alloc(48 * n_out * 4), renderer.alloc(48 * 4)
    bsur.geodesic_knots(renderer, cfg, br, dp, 48, drows, dsteps)
    renderer.synchronize()
    rows = drows.to_host(np.empty((48, n_out), np.float32))
    steps = dsteps.to_host(np.empty(48, np.int32))
    ref, ref_steps = osur.knots(params.astype(np.float64), cfg.cam_r, cfg.step, True, cfg.max_steps,
                                cfg.escape_radius, float(np.float32(br.r[0])), float(np.float32(br.r[-1])), 64)
    K = 64
    # heads and step counts agree on nearly all rays (chaotic rays near the photon ring may differ)
    same = (rows[:, 4 * K] == ref[:, 4 * K]) & (rows[:, 4 * K + 2] == ref[:, 4 * K + 2]) & (steps == ref_steps)
    assert same.mean() > 0.9, same.mean()
    kn, kr = rows[same, :4 * K].astype(np.float64), ref[same, :4 * K]
    assert np.allclose(kn, kr, rtol=2e-6, atol=2e-6 * np.abs(kr).max()), np.abs(kn - kr).max()
    for d in (dp, drows, dsteps):
        d.free()


def test_knot_render_with_exact_knots_matches_rk2(renderer):
    """RK2 knots of every pixel (bh_geodesic_knots) rendered by bh_knot_render_device agree with the FP32
    RK2 render; the fallback rays are traced by the same FP32 kernel and agree bit for bit."""
    print("knot render from exact knots, 64x64 scale scene", flush=True)
    sc, br = scenes.scale_plain_scene(64, 64), scenes.scale_brick()
    cfg = bsur.SurrogateConfig(n_knots=128)
    n = sc.width * sc.height
    params = bsur.pixel_params(sc)
    dp = _upload(renderer, params)
    drows = renderer.alloc(n * cfg.n_out * 4)
    bsur.geodesic_knots(renderer, cfg, br, dp, n, drows)
    renderer.synchronize()
    rows = drows.to_host(np.empty((n, cfg.n_out), np.float32))
    dT = _upload(renderer, np.ascontiguousarray(rows.T))
    db = renderer.upload(br, "f32", "corner")
    ones, zeros = np.ones(cfg.n_out, np.float32), np.zeros(cfg.n_out, np.float32)
    cfg0 = bsur.SurrogateConfig(n_knots=128, crit_distance=0.0, cap_lo=0.5, cap_hi=0.5)
    out = bsur.knot_render(renderer, sc, db, dT, n, cfg0, ones, zeros)
    ref = renderer.render(sc, db, PREC_FP32)
    st = out["status"].ravel()
    sur = (st & 16) != 0
    flagged = rows[:, 4 * 128 + 1] > 0.5
    assert np.array_equal(~sur, flagged)
    assert out["fallback"] == int(flagged.sum())
    I, I0 = out["iquv"].reshape(-1, 4), ref["iquv"].reshape(-1, 4)
    # fallback pixels: the same FP32 RK2 kernel on the same rays
    assert np.array_equal(I[~sur], I0[~sur])
    assert np.array_equal(out["steps"].ravel()[~sur], ref["steps"].ravel()[~sur])
    # surrogate pixels: FP32 transfer over FP64 knots stored as float32
    ok = sur & ((ref["status"].ravel() & 8) == 0)
    e = np.abs(I[ok, 0] - I0[ok, 0]).sum() / np.abs(I0[ok, 0]).sum()
    assert e < 5e-4, e
    assert np.mean((st[ok] & 7) == (ref["status"].ravel()[ok] & 7)) > 0.99
    renderer.release(db)
    for d in (dp, drows, dT):
        d.free()


def test_surrogate_render_gains_over_the_untrained_network(renderer):
    """What the learned part gains over the method without it: after a short training run on seeded RK2
    geodesics the surrogate takes most of the image off the RK2 integrator (the fallback render integrates
    at most half of the plain render's RK2 steps), and its image is far closer to the RK2 image than the
    same pipeline with the untrained network.  Its image error and its wall time are what that costs (the
    MLP and the knot transfer are not cheaper than the RK2 steps they replace; bench.py reports both)."""
    print("surrogate short training + render", flush=True)
    sc, br = scenes.scale_plain_scene(64, 64), scenes.scale_brick()
    cfg = bsur.SurrogateConfig(spin=(0.9, 0.98), inclination_deg=(10.0, 25.0))
    sur = bsur.Surrogate(renderer, br, cfg, seed=1)
    data = sur.generate(1 << 15, 99)
    db = renderer.upload(br, "f32", "corner")
    ref = renderer.render(sc, db, PREC_FP32)
    plain_steps, _ = renderer.last_work()
    out0 = sur.render(sc, db)
    info = sur.train(data, 400, batch=8192, lr=1e-3, log_every=50)
    out1 = sur.render(sc, db)
    fallback_steps, _ = renderer.last_work()
    losses = [l for _, l in sur.history]
    assert losses[-1] < 0.5 * losses[0], losses
    I0 = ref["iquv"][..., 0].astype(np.float64).ravel()
    ok = (ref["status"].ravel() & 8) == 0

    def err(o):
        return np.abs(o["iquv"][..., 0].astype(np.float64).ravel()[ok] - I0[ok]).sum() / np.abs(I0[ok]).sum()

    e0, e1 = err(out0), err(out1)
    print(f"rel L1 on I: untrained {e0:.3e}, trained ({info['steps']} steps) {e1:.3e}, fallback {out1['fallback']}; "
          f"RK2 steps: plain {plain_steps}, surrogate fallback {fallback_steps}")
    assert e1 < 0.5 * e0
    assert 0 < fallback_steps <= 0.5 * plain_steps
    assert out1["fallback"] < sc.width * sc.height
    renderer.release(db)
    for k in ("x", "t", "w"):
        data[k].free()
    sur.close()


@pytest.mark.parametrize("rows", [1, 131, 1000, 4096 + 5])
def test_fused_inference_matches_layer_gemms_and_oracle(renderer, rows):
    """The fused all-layer inference kernel (activations in LDS) against the per-layer GEMM path and
    the BF16-emulating FP64 oracle, for ragged row counts and the surrogate's 515 outputs."""
    print(f"fused inference rows={rows}", flush=True)
    n_out = 515
    mlp = bsur.MLP(renderer, 4, 256, 4, n_out)
    params = _params_for(mlp, 21)
    mlp.set_params(params)
    rng = np.random.default_rng(rows)
    x = rng.uniform(-1, 1, size=(rows, 4)).astype(np.float32)
    dx = _upload(renderer, x)
    dy_f, dy_g = renderer.alloc(rows * n_out * 4), renderer.alloc(rows * n_out * 4)
    mlp.set_fused(True)
    mlp.forward(dx, rows, dy_f, transposed=True)
    mlp.set_fused(False)
    mlp.forward(dx, rows, dy_g, transposed=True)
    yf = dy_f.to_host(np.empty((n_out, rows), np.float32)).astype(np.float64)
    yg = dy_g.to_host(np.empty((n_out, rows), np.float32)).astype(np.float64)
    ref, _ = osur.mlp_forward([(W.astype(np.float64), b.astype(np.float64)) for W, b in params], x, bf16=True)
    ref = ref[:, :n_out].T
    scale = np.abs(ref).max()
    # BF16 hidden activations: an accumulation-order difference can flip one activation's rounding
    # (one BF16 step, 2^-8 relative), so the bound is on the mean error and loose on the worst value
    for y in (yf, yg):
        assert np.abs(y - ref).sum() / np.abs(ref).sum() < 1e-3
        assert np.abs(y - ref).max() / scale < 1e-2
    assert np.abs(yf - yg).sum() / np.abs(yg).sum() < 1e-3
    for d in (dx, dy_f, dy_g):
        d.free()
    mlp.close()


def test_checkpoint_resumes_training_with_the_optimizer_state(renderer, tmp_path):
    """save() keeps Adam's moments and step count; a surrogate loaded from the file takes the same next steps
    as the one that wrote it (to the atomics' summation order), while the same weights with a fresh Adam state
    step elsewhere."""
    print("surrogate checkpoint resume", flush=True)
    br = scenes.scale_brick(32, 16, 32)
    cfg = bsur.SurrogateConfig(spin=(0.9, 0.98), inclination_deg=(10.0, 25.0))
    a = bsur.Surrogate(renderer, br, cfg, seed=3)
    data = a.generate(1 << 13, 5)
    a.train(data, 20, batch=2048, lr=1e-3, log_every=10)
    path = tmp_path / "surrogate.npz"
    a.save(path)
    b = bsur.Surrogate(renderer, br, cfg, seed=4)
    b.load(path)
    ma, va, ta = a.mlp.get_adam()
    mb, vb, tb = b.mlp.get_adam()
    assert ta == tb == 20
    np.testing.assert_array_equal(ma, mb)
    np.testing.assert_array_equal(va, vb)
    c = bsur.Surrogate(renderer, br, cfg, seed=5)
    c.load(path)
    c.mlp.set_params(c.mlp.get_params())   # same weights, Adam state reset
    for s in (a, b, c):
        s.train(data, 5, batch=2048, lr=1e-3, log_every=5)
    pa, pb, pc = (np.concatenate([np.concatenate([W.ravel(), bb]) for W, bb in s.mlp.get_params()]) for s in (a, b, c))
    d_ab, d_ac = float(np.abs(pa - pb).max()), float(np.abs(pa - pc).max())
    print(f"after 5 more steps: max |a - b| {d_ab:.2e} (resumed), max |a - c| {d_ac:.2e} (fresh Adam state)")
    assert d_ab <= 1e-6
    assert d_ac >= 100 * max(d_ab, 1e-9)
    for k in ("x", "t", "w"):
        data[k].free()
    for s in (a, b, c):
        s.close()


def test_surrogate_batch_equals_single_surrogate_renders(renderer):
    """The batched surrogate (inputs of all renders, one FP16 inference, the knot transfer of all tiles, all
    fallback rays in one persistent RK2 launch, captured in a hipGraph and replayed) gives each render what the
    single surrogate render gives: the same fallback pixels, the same surrogate pixels bit for bit (same FP16
    knots, same transfer code), and the same RK2 results on the fallback pixels (statuses and steps; values to
    the lone-ray split's rounding)."""
    print("surrogate batch vs single renders", flush=True)
    bricks = scenes.batch_bricks()[:2]
    cfg = bsur.SurrogateConfig(spin=(0.0, 0.998), inclination_deg=(5.0, 85.0))
    sur = bsur.Surrogate(renderer, bricks[0], cfg, seed=2)
    data = sur.generate(1 << 14, 3)
    sur.train(data, 60, batch=4096, lr=1e-3, log_every=30)
    dbs = [renderer.upload(b, "bf16", "corner") for b in bricks]
    jobs = [(sc.with_size(w, w), dbs[k % 2]) for k, (sc, _), w in zip(range(4), scenes.batch_scenes(4, seed=9),
                                                                       (40, 96, 17, 64))]
    saved = renderer.get_tuning(cfg.fallback_precision)
    renderer.set_tuning(cfg.fallback_precision, coop_steps=0)   # the batched launch does not split lone rays
    sb = sur.batch(jobs)
    try:
        sb.launch(graph=False)
        direct = sb.results()
        per, total = sb.fallback()
        sb.launch(graph=True)
        sb.launch(graph=True)
        replay = sb.results()
        for a, b in zip(direct, replay):
            np.testing.assert_array_equal(a["iquv"], b["iquv"])
            np.testing.assert_array_equal(a["status"], b["status"])
        assert total == sum(per) and 0 < total < sb.pixels
        for (sc, db), got, nfb in zip(jobs, direct, per):
            ref = sur.render(sc, db)
            assert ref["fallback"] == nfb
            sur_px = (ref["status"] & 16) != 0
            np.testing.assert_array_equal((got["status"] & 16) != 0, sur_px)
            np.testing.assert_array_equal(got["iquv"][sur_px], ref["iquv"][sur_px])
            np.testing.assert_array_equal(got["steps"][sur_px], ref["steps"][sur_px])
            fb = ~sur_px
            assert np.mean(got["status"][fb] == ref["status"][fb]) >= 0.99
            assert np.mean(got["steps"][fb] == ref["steps"][fb]) >= 0.99
            I, I0 = got["iquv"][..., 0][fb].astype(np.float64), ref["iquv"][..., 0][fb].astype(np.float64)
            assert np.abs(I - I0).sum() <= 1e-5 * max(np.abs(I0).sum(), 1e-30)
    finally:
        sb.close()
        renderer.set_tuning(cfg.fallback_precision, **saved)
        for d in dbs:
            renderer.release(d)
        for k in ("x", "t", "w"):
            data[k].free()
        sur.close()
