"""Learned surrogate of the Kerr ray tracer: an MLP from (alpha, beta, a, cos i) to geodesic knots.

The network maps a ray's image-plane offsets (alpha, beta, in M), the spin and
the cosine of the inclination to ``n_knots`` knots (r, theta, phi, g) of its RK2
geodesic — the RK2 states themselves, from the start of the first step whose
midpoint radius lies inside the brick — plus three heads: captured (0/1), flag
(the brick part needs more knots than the network has, a pole clamp fired, or
the ray failed) and the number of recorded states.  Training data are FP64 RK2
geodesics traced by ``bh_geodesic_knots``; training runs the library's BF16
MFMA forward/backward and Adam kernels (``bh_mlp_*``).  At render time the
transfer runs over the predicted knots with the RK2 transfer rule; rays within
``crit_distance`` of the critical curve, flagged rays and rays with an
ambiguous captured head are traced by the RK2 kernel instead.

Nothing here runs on the CPU: every step is a library call on the GPU.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import asdict, dataclass

import numpy as np

from . import _lib
from .render import PREC_FP32, DeviceBrick, Renderer, _ptr, scene_desc
from .scenes import Brick, Scene

HEADS = 3


def _off(buf, nbytes: int):
    return C.c_void_p(_ptr(buf).value + int(nbytes))


class MLP:
    """x (n_in) -> [Linear width, SiLU] x depth -> Linear n_out on the GPU (bh_mlp_*)."""

    def __init__(self, renderer: Renderer, n_in: int, width: int, depth: int, n_out: int):
        self.r = renderer
        self._lib = renderer._lib
        self.n_in, self.width, self.depth, self.n_out = int(n_in), int(width), int(depth), int(n_out)
        h = C.c_void_p()
        _lib.check(self._lib.bh_mlp_create(renderer.ctx, self.n_in, self.width, self.depth, self.n_out, C.byref(h)))
        self.h = h
        a, b, n = C.c_int32(), C.c_int32(), C.c_int64()
        _lib.check(self._lib.bh_mlp_shape(h, C.byref(a), C.byref(b), C.byref(n)))
        dims = [int(a.value)] + [self.width] * self.depth + [int(b.value)]
        self.shapes = [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]
        self.n_params = int(n.value)
        renderer.adopt(self)

    def close(self) -> None:
        if self.h and self.r.ctx:
            self._lib.bh_mlp_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters: list of (W [fi_pad, fo_pad], b [fo_pad]) float32
    def init_params(self, seed: int = 0) -> list:
        """He-uniform weights (fan-in of the unpadded inputs), zero biases; padded entries zero."""
        rng = np.random.default_rng(int(seed))
        out = []
        for l, (fi, fo) in enumerate(self.shapes):
            real_in = self.n_in if l == 0 else fi
            real_out = self.n_out if l == len(self.shapes) - 1 else fo
            lim = math.sqrt(6.0 / real_in) if l < len(self.shapes) - 1 else math.sqrt(3.0 / real_in)
            W = np.zeros((fi, fo), dtype=np.float32)
            W[:real_in, :real_out] = rng.uniform(-lim, lim, size=(real_in, real_out))
            out.append((W, np.zeros(fo, dtype=np.float32)))
        return out

    def flat(self, params) -> np.ndarray:
        return np.concatenate([np.concatenate([np.asarray(W, np.float32).ravel(), np.asarray(b, np.float32).ravel()])
                               for W, b in params]).astype(np.float32)

    def unflat(self, vec) -> list:
        out, off = [], 0
        for fi, fo in self.shapes:
            W = np.asarray(vec[off:off + fi * fo], np.float32).reshape(fi, fo)
            off += fi * fo
            out.append((W, np.asarray(vec[off:off + fo], np.float32)))
            off += fo
        return out

    def set_params(self, params) -> None:
        v = np.ascontiguousarray(self.flat(params))
        assert v.size == self.n_params
        _lib.check(self._lib.bh_mlp_set_params(self.h, _lib.fptr(v)))

    def get_params(self) -> list:
        v = np.empty(self.n_params, dtype=np.float32)
        _lib.check(self._lib.bh_mlp_get_params(self.h, _lib.fptr(v)))
        return self.unflat(v)

    def get_adam(self) -> tuple[np.ndarray, np.ndarray, int]:
        """Adam's first and second moments (flat, the parameter layout) and its step count."""
        m, v, t = np.empty(self.n_params, np.float32), np.empty(self.n_params, np.float32), C.c_int64(0)
        _lib.check(self._lib.bh_mlp_get_adam(self.h, _lib.fptr(m), _lib.fptr(v), C.byref(t)))
        return m, v, int(t.value)

    def set_adam(self, m: np.ndarray, v: np.ndarray, steps: int) -> None:
        m, v = (np.ascontiguousarray(a, dtype=np.float32) for a in (m, v))
        assert m.size == self.n_params and v.size == self.n_params
        _lib.check(self._lib.bh_mlp_set_adam(self.h, _lib.fptr(m), _lib.fptr(v), int(steps)))

    def set_fused(self, fused: bool) -> None:
        """Transposed inference through the fused all-layer kernel (default) or the per-layer GEMMs."""
        _lib.check(self._lib.bh_mlp_set_fused(self.h, int(bool(fused))))

    def forward(self, d_x, rows: int, d_y, transposed: bool = False) -> None:
        _lib.check(self._lib.bh_mlp_forward(self.h, _ptr(d_x), int(rows), _ptr(d_y), int(bool(transposed))))

    def gradient(self, d_x, d_t, d_w, rows: int, d_row_w=None, masked_cols: int = 0) -> tuple[float, list]:
        """Loss and dLoss/dparams; d_row_w [rows] weights the first masked_cols outputs of each row."""
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_float(0.0)
        _lib.check(self._lib.bh_mlp_gradient(self.h, _ptr(d_x), _ptr(d_t), _ptr(d_w), _ptr(d_row_w), int(masked_cols),
                                             int(rows), _lib.fptr(g), C.byref(loss)))
        return float(loss.value), self.unflat(g)

    def train_step(self, d_x, d_t, d_w, rows: int, lr: float, want_loss: bool = False, d_row_w=None,
                   masked_cols: int = 0):
        loss = C.c_float(0.0)
        _lib.check(self._lib.bh_mlp_train_step(self.h, _ptr(d_x), _ptr(d_t), _ptr(d_w), _ptr(d_row_w),
                                               int(masked_cols), int(rows), float(lr),
                                               C.byref(loss) if want_loss else None))
        return float(loss.value) if want_loss else None


@dataclass
class SurrogateConfig:
    """Network shape, training distribution, geodesic rule and fallback thresholds."""
    n_knots: int = 128
    width: int = 256
    depth: int = 4
    cam_r: float = 1000.0
    half_width: float = 20.0                 # alpha, beta ~ U[-half_width, half_width] M
    spin: tuple = (0.0, 0.998)               # a ~ U
    inclination_deg: tuple = (5.0, 85.0)     # i ~ U (degrees)
    step: float = 0.05                       # h = step (r - r_+)
    step_mode: int = 1
    max_steps: int = 4096
    escape_radius: float = 220.0
    crit_distance: float = 0.5               # M, Bardeen image plane
    flag_max: float = 0.5
    cap_lo: float = 0.25
    cap_hi: float = 0.75
    fallback_precision: int = PREC_FP32
    head_weight: float = 1.0

    @property
    def n_out(self) -> int:
        return 4 * self.n_knots + HEADS


def sample_params(n: int, seed: int, cfg: SurrogateConfig) -> np.ndarray:
    """Seeded rays (n, 4) float32: alpha, beta, a, cos i."""
    rng = np.random.default_rng(int(seed))
    p = np.empty((int(n), 4), dtype=np.float32)
    p[:, 0] = rng.uniform(-cfg.half_width, cfg.half_width, n)
    p[:, 1] = rng.uniform(-cfg.half_width, cfg.half_width, n)
    p[:, 2] = rng.uniform(cfg.spin[0], cfg.spin[1], n)
    p[:, 3] = np.cos(np.radians(rng.uniform(cfg.inclination_deg[0], cfg.inclination_deg[1], n)))
    return p


def pixel_params(scene: Scene) -> np.ndarray:
    """(alpha, beta, a, cos i) of every pixel of a render, as the library's surrogate input kernel forms them."""
    cam = scene.camera
    W, H = int(scene.width), int(scene.height)
    px, py = np.meshgrid(np.arange(W), np.arange(H))
    ndc_x = 2.0 * (px.ravel() + 0.5) / W - 1.0
    ndc_y = 1.0 - 2.0 * (py.ravel() + 0.5) / H
    tan_y = math.tan(0.5 * (cam.fov_y_degrees * (math.pi / 180.0)))
    tan_x = (W / H) * tan_y
    n_th, n_ph = -ndc_y * tan_y, ndc_x * tan_x
    if cam.roll_degrees:
        ro = cam.roll_degrees * (math.pi / 180.0)
        n_th, n_ph = math.cos(ro) * n_th - math.sin(ro) * n_ph, math.sin(ro) * n_th + math.cos(ro) * n_ph
    out = np.empty((W * H, 4), dtype=np.float32)
    out[:, 0] = (n_ph * cam.r).astype(np.float32)
    out[:, 1] = (-n_th * cam.r).astype(np.float32)
    out[:, 2] = np.float32(scene.spin)
    out[:, 3] = np.float32(math.cos(min(max(cam.theta, 1e-7), math.pi - 1e-7)))
    return out


def knot_desc(cfg: SurrogateConfig, brick: Brick) -> _lib.KnotDesc:
    return _lib.KnotDesc(cam_r=cfg.cam_r, step=cfg.step, escape_radius=cfg.escape_radius,
                         r_lo=float(np.float32(brick.r[0])), r_hi=float(np.float32(brick.r[-1])),
                         step_mode=cfg.step_mode, max_steps=cfg.max_steps, n_knots=cfg.n_knots, reserved=0)


def geodesic_knots(renderer: Renderer, cfg: SurrogateConfig, brick: Brick, d_params, n: int, d_rows,
                   d_steps=None, d_weight=None) -> None:
    """Enqueue bh_geodesic_knots: rows (n, 4 * n_knots + 3) float32 for rays d_params (n, 4)."""
    desc = knot_desc(cfg, brick)
    _lib.check(renderer._lib.bh_geodesic_knots(renderer.ctx, C.byref(desc), _ptr(d_params), int(n), _ptr(d_rows),
                                               _ptr(d_steps), _ptr(d_weight)))


@dataclass
class Normalisation:
    in_mul: np.ndarray
    in_add: np.ndarray
    out_mul: np.ndarray    # raw = y * out_mul + out_add
    out_add: np.ndarray


class Surrogate:
    """MLP surrogate bound to one Renderer context and one brick's radial range."""

    def __init__(self, renderer: Renderer, brick: Brick, cfg: SurrogateConfig | None = None, seed: int = 0):
        self.r = renderer
        self.cfg = cfg or SurrogateConfig()
        self.brick = brick
        self.mlp = MLP(renderer, 4, self.cfg.width, self.cfg.depth, self.cfg.n_out)
        self.mlp.set_params(self.mlp.init_params(seed))
        self.norm: Normalisation | None = None
        self.history: list = []
        self._work = None
        self._work_rows = -1
        renderer.adopt(self)

    # ---- training data -----------------------------------------------------------
    def input_normalisation(self) -> tuple[np.ndarray, np.ndarray]:
        c = self.cfg
        mul = np.array([1.0 / c.half_width, 1.0 / c.half_width, 2.0 / max(c.spin[1] - c.spin[0], 1e-3), 2.0],
                       dtype=np.float32)
        ci = np.cos(np.radians(np.array(c.inclination_deg, dtype=np.float64)))
        add = np.array([0.0, 0.0, 0.5 * (c.spin[0] + c.spin[1]), 0.5 * (ci[0] + ci[1])], dtype=np.float32)
        mul[3] = np.float32(2.0 / max(abs(ci[0] - ci[1]), 1e-3))
        return mul, add

    def generate(self, n: int, seed: int, stats_rows: int = 65536) -> dict:
        """Trace n seeded rays into device buffers; normalise inputs and targets in place.

        Returns dict(x=DeviceBuffer (n, 4), t=DeviceBuffer (n, n_out), w=DeviceBuffer (n,) ray weights
        (0 for flagged rays, whose knots are left out of the loss), steps=int32 host, flagged=fraction,
        ms=kernel time)."""
        c = self.cfg
        p = sample_params(n, seed, c)
        d_p = self.r.alloc(p.nbytes)
        d_p.from_host(p)
        d_t = self.r.alloc(n * c.n_out * 4)
        d_s = self.r.alloc(n * 4)
        d_w = self.r.alloc(n * 4)
        geodesic_knots(self.r, c, self.brick, d_p, n, d_t, d_s, d_w)
        self.r.synchronize()
        ms = self.r.last_kernel_ms()
        steps = np.empty(n, dtype=np.int32)
        d_s.to_host(steps)
        d_s.free()
        w = d_w.to_host(np.empty(n, dtype=np.float32))
        if self.norm is None:
            m = min(int(stats_rows), n)
            sub = np.empty((m, c.n_out), dtype=np.float32)
            d_t.to_host(sub)
            good = sub[w[:m] > 0.5].astype(np.float64)
            if good.shape[0] < 2:
                good = sub.astype(np.float64)
            mean = good.mean(axis=0)
            std = good.std(axis=0)
            # heads keep their 0/1 (and state-count) scale over all rays
            heads = slice(4 * c.n_knots, c.n_out)
            mean[heads] = sub[:, heads].astype(np.float64).mean(axis=0)
            std[heads] = sub[:, heads].astype(np.float64).std(axis=0)
            std = np.where(std > 1e-6, std, 1.0)
            in_mul, in_add = self.input_normalisation()
            self.norm = Normalisation(in_mul, in_add, std.astype(np.float32), mean.astype(np.float32))
        nm = self.norm
        # targets: (raw - add) / mul; inputs: (p - add) * mul
        _lib.check(self.r._lib.bh_columns_affine(self.r.ctx, _ptr(d_t), n, c.n_out,
                                                 _lib.fptr(np.ascontiguousarray(1.0 / nm.out_mul, np.float32)),
                                                 _lib.fptr(np.ascontiguousarray(-nm.out_add / nm.out_mul, np.float32))))
        _lib.check(self.r._lib.bh_columns_affine(self.r.ctx, _ptr(d_p), n, 4,
                                                 _lib.fptr(np.ascontiguousarray(nm.in_mul, np.float32)),
                                                 _lib.fptr(np.ascontiguousarray(-nm.in_add * nm.in_mul, np.float32))))
        return {"x": d_p, "t": d_t, "w": d_w, "steps": steps, "ms": ms, "n": n, "flagged": float(1.0 - w.mean())}

    def loss_weights(self) -> np.ndarray:
        w = np.ones(self.cfg.n_out, dtype=np.float32)
        w[4 * self.cfg.n_knots:] = self.cfg.head_weight
        return w

    # ---- training ----------------------------------------------------------------
    def train(self, data: dict, steps: int, batch: int = 8192, lr: float = 1e-3, log_every: int = 100,
              time_limit_s: float | None = None, first_batch: int = 0) -> dict:
        """Adam on consecutive batches of the (iid, seeded) training rays, from batch ``first_batch`` of the
        epoch (so that successive calls can continue through it)."""
        c = self.cfg
        n = int(data["n"])
        nb = max(n // int(batch), 1)
        w = self.loss_weights()
        d_w = self.r.alloc(w.nbytes)
        d_w.from_host(w)
        t0 = time.perf_counter()
        done = 0
        for s in range(int(steps)):
            b = (int(first_batch) + s) % nb
            want = (s + 1) % int(log_every) == 0 or s == int(steps) - 1
            loss = self.mlp.train_step(_off(data["x"], b * batch * 16), _off(data["t"], b * batch * c.n_out * 4), d_w,
                                       batch, lr, want_loss=want, d_row_w=_off(data["w"], b * batch * 4),
                                       masked_cols=4 * c.n_knots)
            done = s + 1
            if want:
                self.history.append((done, loss))
                if time_limit_s is not None and time.perf_counter() - t0 > time_limit_s:
                    break
        self.r.synchronize()
        d_w.free()
        return {"steps": done, "seconds": time.perf_counter() - t0, "loss": self.history[-1][1] if self.history else None}

    def evaluate(self, data: dict, rows: int | None = None) -> float:
        """Weighted loss of the network on (normalised) data rows, without an update."""
        rows = int(rows or data["n"])
        d_w = self.r.alloc(self.cfg.n_out * 4)
        d_w.from_host(self.loss_weights())
        loss, _ = self.mlp.gradient(data["x"], data["t"], d_w, rows, data["w"], 4 * self.cfg.n_knots)
        d_w.free()
        return loss

    # ---- rendering -----------------------------------------------------------------
    def desc(self) -> _lib.SurrogateDesc:
        nm = self.norm
        if nm is None:
            raise RuntimeError("the surrogate has no normalisation: generate training data or load a checkpoint first")
        self._mul = np.ascontiguousarray(nm.out_mul, np.float32)
        self._add = np.ascontiguousarray(nm.out_add, np.float32)
        return _desc(self.cfg, self._mul, self._add, nm.in_mul, nm.in_add)

    def work_bytes(self, rows: int) -> int:
        b = C.c_size_t(0)
        _lib.check(self.r._lib.bh_surrogate_work_bytes(int(rows), self.cfg.n_knots, C.byref(b)))
        return int(b.value)

    def render_device(self, scene: Scene, dbrick: DeviceBrick, d_iquv, d_steps=None, d_status=None) -> None:
        """Enqueue a surrogate render into device buffers (FP32 IQUV)."""
        if abs(float(scene.camera.r) - self.cfg.cam_r) > 1e-9 * self.cfg.cam_r:
            raise ValueError(f"the surrogate was trained for a camera at r = {self.cfg.cam_r}")
        rows = int(scene.width) * int(scene.height)
        if self._work is None or self._work_rows < rows:
            if self._work is not None:
                self._work.free()
            self._work = self.r.alloc(self.work_bytes(rows))
            self._work_rows = rows
        desc = scene_desc(scene)
        sd = self.desc()
        _lib.check(self.r._lib.bh_surrogate_render_device(self.r.ctx, C.byref(desc), dbrick.brick_id, self.mlp.h,
                                                          C.byref(sd), _ptr(self._work), self._work.nbytes,
                                                          _ptr(d_iquv), _ptr(d_steps), _ptr(d_status)))

    def render(self, scene: Scene, dbrick: DeviceBrick) -> dict:
        """Synchronous surrogate render to host arrays (iquv float32 (H, W, 4), steps, status)."""
        H, W = int(scene.height), int(scene.width)
        n = H * W
        bi, bs, bt = self.r.alloc(n * 16), self.r.alloc(n * 4), self.r.alloc(n)
        try:
            self.render_device(scene, dbrick, bi, bs, bt)
            self.r.synchronize()
            out = {"iquv": bi.to_host(np.empty((H, W, 4), np.float32)), "steps": bs.to_host(np.empty((H, W), np.int32)),
                   "status": bt.to_host(np.empty((H, W), np.uint8)), "ms": self.r.last_kernel_ms(),
                   "fallback": self.last_fallback()}
        finally:
            bi.free()
            bs.free()
            bt.free()
        return out

    def last_fallback(self) -> int:
        v = C.c_int64(0)
        _lib.check(self.r._lib.bh_surrogate_last_fallback(self.r.ctx, C.byref(v)))
        return int(v.value)

    def batch(self, jobs) -> "SurrogateBatch":
        """Many renders through the surrogate in one captured sequence (bh_surrogate_batch_*)."""
        return SurrogateBatch(self, jobs)

    # ---- checkpoints ---------------------------------------------------------------------
    def save(self, path) -> None:
        """Weights, normalisation, loss history and the optimizer state (Adam's moments and step count)."""
        params = self.mlp.get_params()
        arrs = {f"W{i}": W for i, (W, _) in enumerate(params)}
        arrs.update({f"b{i}": b for i, (_, b) in enumerate(params)})
        adam_m, adam_v, adam_t = self.mlp.get_adam()
        nm = self.norm
        np.savez(path, in_mul=nm.in_mul, in_add=nm.in_add, out_mul=nm.out_mul, out_add=nm.out_add,
                 history=np.array(self.history, dtype=np.float64).reshape(-1, 2), adam_m=adam_m, adam_v=adam_v,
                 adam_t=np.array(adam_t, dtype=np.int64), **arrs)

    def load(self, path) -> None:
        """Restore a checkpoint of ``save``; training continues from its Adam state when the file has one."""
        with np.load(path, allow_pickle=False) as z:
            params = [(z[f"W{i}"], z[f"b{i}"]) for i in range(len(self.mlp.shapes))]
            self.norm = Normalisation(z["in_mul"], z["in_add"], z["out_mul"], z["out_add"])
            adam = (z["adam_m"], z["adam_v"], int(z["adam_t"])) if "adam_m" in z.files else None
            self.history = [tuple(h) for h in z["history"]] if "history" in z.files else []
        self.mlp.set_params(params)
        if adam is not None:
            self.mlp.set_adam(*adam)

    def close(self) -> None:
        if self._work is not None and self.r.ctx:
            self._work.free()
        self._work = None
        self.mlp.close()


class SurrogateBatch:
    """Renders (Scene, DeviceBrick) through the surrogate in one sequence: inputs of every render's rays,
    one FP16 fused inference, the knot transfer of every tile, and all fallback rays in one persistent
    RK2 launch; captured in a hipGraph on the first ``launch(graph=True)`` and replayed afterwards."""

    def __init__(self, sur: Surrogate, jobs):
        self.sur = sur
        self.r = sur.r
        self.scenes = [sc for sc, _ in jobs]
        for sc in self.scenes:
            if abs(float(sc.camera.r) - sur.cfg.cam_r) > 1e-9 * sur.cfg.cam_r:
                raise ValueError(f"the surrogate was trained for a camera at r = {sur.cfg.cam_r}")
        n = len(self.scenes)
        items = (_lib.BatchItem * n)()
        for i, (sc, db) in enumerate(jobs):
            items[i].scene = scene_desc(sc)
            items[i].brick_id = db.brick_id
        off = (C.c_int64 * (n + 1))()
        sd = sur.desc()
        h = C.c_void_p()
        _lib.check(self.r._lib.bh_surrogate_batch_create(self.r.ctx, items, n, sur.mlp.h, C.byref(sd), off, C.byref(h)))
        self.handle = h
        self.offsets = [int(off[i]) for i in range(n + 1)]
        self.pixels = self.offsets[-1]
        self._bufs = None

    def buffers(self):
        if self._bufs is None:
            self._bufs = (self.r.alloc(self.pixels * 16), self.r.alloc(self.pixels * 4), self.r.alloc(self.pixels))
        return self._bufs

    def launch(self, graph: bool = True) -> None:
        d_iquv, d_steps, d_status = self.buffers()
        _lib.check(self.r._lib.bh_surrogate_batch_launch(self.handle, d_iquv.ptr, d_steps.ptr, d_status.ptr,
                                                         int(bool(graph))))

    def fallback(self) -> tuple[list, int]:
        """Fallback rays of the last launch: per render, and in all."""
        n = len(self.scenes)
        per = (C.c_int64 * n)()
        tot = C.c_int64(0)
        _lib.check(self.r._lib.bh_surrogate_batch_fallback(self.handle, per, C.byref(tot)))
        return [int(per[i]) for i in range(n)], int(tot.value)

    def results(self) -> list[dict]:
        """Synchronise and return per-render dicts of iquv (H, W, 4) float32, steps and status."""
        d_iquv, d_steps, d_status = self.buffers()
        iquv = d_iquv.to_host(np.empty((self.pixels, 4), dtype=np.float32))
        steps = d_steps.to_host(np.empty(self.pixels, dtype=np.int32))
        status = d_status.to_host(np.empty(self.pixels, dtype=np.uint8))
        out = []
        for i, sc in enumerate(self.scenes):
            a, b = self.offsets[i], self.offsets[i + 1]
            H, W = sc.height, sc.width
            out.append({"iquv": iquv[a:b].reshape(H, W, 4), "steps": steps[a:b].reshape(H, W),
                        "status": status[a:b].reshape(H, W)})
        return out

    def close(self) -> None:
        if self._bufs is not None:
            for b in self._bufs:
                b.free()
            self._bufs = None
        if self.handle and self.r.ctx:
            self.r._lib.bh_surrogate_batch_destroy(self.handle)
        self.handle = None


def _desc(cfg: SurrogateConfig, out_mul: np.ndarray, out_add: np.ndarray, in_mul=None, in_add=None):
    d = _lib.SurrogateDesc()
    for i in range(4):
        d.in_mul[i] = float(in_mul[i]) if in_mul is not None else 1.0
        d.in_add[i] = float(in_add[i]) if in_add is not None else 0.0
    d.out_mul = _lib.fptr(out_mul)
    d.out_add = _lib.fptr(out_add)
    d.n_knots = cfg.n_knots
    d.fallback_precision = int(cfg.fallback_precision)
    d.crit_distance = float(cfg.crit_distance)
    d.flag_max, d.cap_lo, d.cap_hi = float(cfg.flag_max), float(cfg.cap_lo), float(cfg.cap_hi)
    return d


def knot_render(renderer: Renderer, scene: Scene, dbrick: DeviceBrick, d_rows, ld: int, cfg: SurrogateConfig,
                out_mul: np.ndarray, out_add: np.ndarray) -> dict:
    """bh_knot_render_device on caller-supplied knot rows ([n_out][ld] float32), synchronous, host outputs."""
    H, W = int(scene.height), int(scene.width)
    n = H * W
    need = C.c_size_t(0)
    _lib.check(renderer._lib.bh_surrogate_work_bytes(0, cfg.n_knots, C.byref(need)))
    work = renderer.alloc(int(need.value) + n * 4 + 256)
    bi, bs, bt = renderer.alloc(n * 16), renderer.alloc(n * 4), renderer.alloc(n)
    mul = np.ascontiguousarray(out_mul, np.float32)
    add = np.ascontiguousarray(out_add, np.float32)
    try:
        desc, sd = scene_desc(scene), _desc(cfg, mul, add)
        _lib.check(renderer._lib.bh_knot_render_device(renderer.ctx, C.byref(desc), dbrick.brick_id, _ptr(d_rows),
                                                       int(ld), C.byref(sd), _ptr(work), work.nbytes, _ptr(bi),
                                                       _ptr(bs), _ptr(bt)))
        renderer.synchronize()
        v = C.c_int64(0)
        _lib.check(renderer._lib.bh_surrogate_last_fallback(renderer.ctx, C.byref(v)))
        out = {"iquv": bi.to_host(np.empty((H, W, 4), np.float32)), "steps": bs.to_host(np.empty((H, W), np.int32)),
               "status": bt.to_host(np.empty((H, W), np.uint8)), "ms": renderer.last_kernel_ms(),
               "fallback": int(v.value)}
    finally:
        for b in (work, bi, bs, bt):
            b.free()
    return out


def train_surrogate(renderer: Renderer, brick: Brick, cfg: SurrogateConfig | None = None, n_rays: int = 1 << 20,
                    seed: int = 20240605, epochs: float = 8.0, batch: int = 8192, lr: float = 1e-3,
                    time_limit_s: float | None = None, log_every: int = 128) -> tuple[Surrogate, dict]:
    """Generate n_rays seeded RK2 geodesics on the GPU and train the surrogate on them."""
    sur = Surrogate(renderer, brick, cfg, seed=seed)
    t0 = time.perf_counter()
    data = sur.generate(n_rays, seed)
    t_gen = time.perf_counter() - t0
    steps = int(round(epochs * max(n_rays // batch, 1)))
    info = sur.train(data, steps, batch=batch, lr=lr, time_limit_s=time_limit_s, log_every=log_every)
    info.update({"n_rays": n_rays, "generate_s": t_gen, "knots_kernel_ms": data["ms"],
                 "flagged": None, "config": asdict(sur.cfg)})
    info["flagged"] = data["flagged"]
    for k in ("x", "t", "w"):
        data[k].free()
    return sur, info
