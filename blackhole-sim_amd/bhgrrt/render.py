"""Render API over the fused HIP kernels.

``Renderer`` owns one device context (one HIP stream).  ``render`` mirrors
``render_stokes_image_bricks`` (python/blackhole_sim/accelerated_renderer.py:153-158):
one Kerr ray per pixel, RK2 geodesic, brick transfer, image (H, W, 4).
"""

from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib
from .scenes import Brick, Scene

PREC_FP64, PREC_FP32, PREC_MIXED = 0, 1, 2
STATUS_NAMES = {0: "max_steps", 1: "captured", 2: "escaped", 3: "invalid", 4: "opaque"}
POLE_CLAMPED = 8


def scene_desc(scene: Scene) -> _lib.SceneDesc:
    cam = scene.camera
    return _lib.SceneDesc(
        spin=float(scene.spin), cam_r=float(cam.r), cam_theta=float(cam.theta), cam_phi=float(cam.phi),
        fov_y_degrees=float(cam.fov_y_degrees), roll_degrees=float(cam.roll_degrees),
        width=int(scene.width), height=int(scene.height), step_mode=int(scene.step_mode),
        max_steps=int(scene.max_steps), step=float(scene.step), escape_radius=float(scene.escape_radius),
        affine_to_length_cm=float(scene.affine_to_length_cm), max_optical_depth=float(scene.max_optical_depth),
        transfer=_lib.TRANSFER_EXACT if scene.exact_step else _lib.TRANSFER_RK2,
        integrator={"rk2": _lib.INTEGRATOR_RK2, "rk4": _lib.INTEGRATOR_RK4}[scene.integrator],
        sampling=_lib.SAMPLING_NUMPY32 if scene.numpy32_sampling else _lib.SAMPLING_NATIVE)


@dataclass
class DeviceBrick:
    brick_id: int
    storage: str
    shape: tuple
    layout: str = "soa"


class Renderer:
    """One GPU context: uploaded bricks, tuning and the render entry points."""

    def __init__(self, device: int = 0):
        self._lib = _lib.lib()
        h = C.c_void_p()
        _lib.check(self._lib.bh_ctx_create(int(device), C.byref(h)))
        self.ctx = h
        self.device = int(device)
        self._dependents = weakref.WeakSet()   # objects holding library handles on this context

    def adopt(self, obj) -> None:
        """Register an object whose ``close()`` must run before the context is destroyed."""
        self._dependents.add(obj)

    def close(self) -> None:
        if self.ctx:
            for obj in list(self._dependents):
                obj.close()
            self._lib.bh_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- bricks -------------------------------------------------------------
    def upload(self, brick: Brick, storage: str = "f32", layout: str = "soa") -> DeviceBrick:
        """Upload an AoS brick; storage f32/bf16/f16, layout soa (channel planes) or corner (per-cell records)."""
        c = np.ascontiguousarray(brick.coeffs, dtype=np.float32)
        bid = C.c_int32(-1)
        _lib.check(self._lib.bh_brick_upload(self.ctx, _lib.fptr(c), _lib.fptr(brick.r), brick.r.size,
                                             _lib.fptr(brick.theta), brick.theta.size, _lib.fptr(brick.phi),
                                             brick.phi.size, _lib.STORE[storage], _lib.LAYOUT[layout], C.byref(bid)))
        return DeviceBrick(int(bid.value), storage, tuple(c.shape), layout)

    def release(self, dbrick: DeviceBrick) -> None:
        _lib.check(self._lib.bh_brick_release(self.ctx, dbrick.brick_id))

    # --- tuning -------------------------------------------------------------
    def set_tuning(self, precision: int, **kw) -> None:
        t = _lib.Tuning(block_threads=kw.get("block_threads", 0), blocks_per_cu=kw.get("blocks_per_cu", -1),
                        steps_per_sync=kw.get("steps_per_sync", 0), refill_min=kw.get("refill_min", 0),
                        queues=kw.get("queues", 0), order=kw.get("order", -1), coop_steps=kw.get("coop_steps", -1))
        _lib.check(self._lib.bh_ctx_set_tuning(self.ctx, int(precision), C.byref(t)))

    def get_tuning(self, precision: int) -> dict:
        t = _lib.Tuning()
        _lib.check(self._lib.bh_ctx_get_tuning(self.ctx, int(precision), C.byref(t)))
        return {f: int(getattr(t, f)) for f, _ in _lib.Tuning._fields_}

    # --- render -------------------------------------------------------------
    def render(self, scene: Scene, dbrick: DeviceBrick, precision: int = PREC_FP32) -> dict:
        """Synchronous render to host arrays: iquv (H, W, 4), steps (H, W) int32, status (H, W) uint8."""
        H, W = int(scene.height), int(scene.width)
        iquv = np.empty((H, W, 4), dtype=np.float64 if precision == PREC_FP64 else np.float32)
        steps = np.empty((H, W), dtype=np.int32)
        status = np.empty((H, W), dtype=np.uint8)
        desc = scene_desc(scene)
        _lib.check(self._lib.bh_render(self.ctx, C.byref(desc), dbrick.brick_id, int(precision),
                                       iquv.ctypes.data_as(C.c_void_p), steps.ctypes.data_as(_lib._i32p),
                                       status.ctypes.data_as(_lib._u8p)))
        return {"iquv": iquv, "steps": steps, "status": status}

    def render_device(self, scene: Scene, dbrick: DeviceBrick, precision: int, d_iquv, d_steps=None,
                      d_status=None) -> None:
        """Enqueue a render into device buffers (``DeviceBuffer`` or raw pointers); no synchronisation."""
        desc = scene_desc(scene)
        _lib.check(self._lib.bh_render_device(self.ctx, C.byref(desc), dbrick.brick_id, int(precision),
                                              _ptr(d_iquv), _ptr(d_steps), _ptr(d_status)))

    def render_device_noalloc(self, scene: Scene, dbrick: DeviceBrick, precision: int, d_iquv, d_steps=None,
                              d_status=None) -> None:
        """``render_device`` without device allocation: counters and tile schedule in the library's static
        per-GPU array (images up to 2048 x 2048); calls overlapping on one GPU race."""
        desc = scene_desc(scene)
        _lib.check(self._lib.bh_render_device_noalloc(self.ctx, C.byref(desc), dbrick.brick_id, int(precision),
                                                      _ptr(d_iquv), _ptr(d_steps), _ptr(d_status)))

    def synchronize(self) -> None:
        _lib.check(self._lib.bh_ctx_synchronize(self.ctx))

    def last_kernel_ms(self) -> float:
        v = C.c_double(0.0)
        _lib.check(self._lib.bh_ctx_last_kernel_ms(self.ctx, C.byref(v)))
        return float(v.value)

    def last_work(self) -> tuple[int, int]:
        """(RK2 steps, in-brick samples) of the last render launch."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        _lib.check(self._lib.bh_ctx_last_work(self.ctx, C.byref(a), C.byref(b), None))
        return int(a.value), int(b.value)

    def last_coop(self) -> tuple[int, int]:
        """(lone rays split over two waves, hand-overs that timed out) of the last render."""
        a, b = C.c_uint64(), C.c_uint64()
        _lib.check(self._lib.bh_ctx_last_coop(self.ctx, C.byref(a), C.byref(b)))
        return int(a.value), int(b.value)

    def last_lane_utilisation(self) -> float:
        """RK2 steps / lane slots of the last launch: SIMD lane occupancy after termination and compaction."""
        a, c = C.c_uint64(0), C.c_uint64(0)
        _lib.check(self._lib.bh_ctx_last_work(self.ctx, C.byref(a), None, C.byref(c)))
        return float(a.value) / max(float(c.value), 1.0)

    # --- device memory ----------------------------------------------------------
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)


class DeviceBuffer:
    """Raw device allocation owned by a Renderer context."""

    def __init__(self, renderer: Renderer, nbytes: int):
        self.r = renderer
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _lib.check(renderer._lib.bh_device_alloc(renderer.ctx, self.nbytes, C.byref(p)))
        self.ptr = p

    def to_host(self, out: np.ndarray) -> np.ndarray:
        assert out.nbytes <= self.nbytes
        _lib.check(self.r._lib.bh_copy_to_host(self.r.ctx, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def from_host(self, src: np.ndarray) -> None:
        src = np.ascontiguousarray(src)
        assert src.nbytes <= self.nbytes
        _lib.check(self.r._lib.bh_copy_to_device(self.r.ctx, self.ptr, src.ctypes.data_as(C.c_void_p), src.nbytes))

    def free(self) -> None:
        if self.ptr and self.r.ctx:
            _lib.check(self.r._lib.bh_device_free(self.r.ctx, self.ptr))
        self.ptr = None


class Batch:
    """Many renders traced by one persistent launch (bh_batch_*); outputs are concatenated per pixel."""

    def __init__(self, renderer: Renderer, jobs, precision: int = PREC_FP32):
        """jobs: sequence of (Scene, DeviceBrick); all bricks must share storage and layout."""
        self.r = renderer
        self.scenes = [sc for sc, _ in jobs]
        self.precision = int(precision)
        n = len(self.scenes)
        items = (_lib.BatchItem * n)()
        for i, (sc, db) in enumerate(jobs):
            items[i].scene = scene_desc(sc)
            items[i].brick_id = db.brick_id
        off = (C.c_int64 * (n + 1))()
        h = C.c_void_p()
        _lib.check(renderer._lib.bh_batch_create(renderer.ctx, items, n, self.precision, off, C.byref(h)))
        self.handle = h
        self.offsets = [int(off[i]) for i in range(n + 1)]
        self.pixels = self.offsets[-1]
        self._bufs = None
        renderer.adopt(self)

    def buffers(self):
        if self._bufs is None:
            esz = 8 if self.precision == PREC_FP64 else 4
            self._bufs = (self.r.alloc(self.pixels * 4 * esz), self.r.alloc(self.pixels * 4), self.r.alloc(self.pixels))
        return self._bufs

    def launch(self, graph: bool = True) -> None:
        """Enqueue one batched launch into the batch's own device outputs."""
        d_iquv, d_steps, d_status = self.buffers()
        fn = self.r._lib.bh_batch_launch_graph if graph else self.r._lib.bh_batch_launch
        _lib.check(fn(self.handle, d_iquv.ptr, d_steps.ptr, d_status.ptr))

    def results(self) -> list[dict]:
        """Synchronise and return per-render dicts of iquv (H, W, 4), steps and status."""
        d_iquv, d_steps, d_status = self.buffers()
        esz = np.float64 if self.precision == PREC_FP64 else np.float32
        iquv = d_iquv.to_host(np.empty((self.pixels, 4), dtype=esz))
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
            self.r._lib.bh_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return x.ptr
    return x


def render_stokes_image_bricks(scene: Scene, brick: Brick, precision: int = PREC_FP32,
                               renderer: Renderer | None = None) -> np.ndarray:
    """accelerated_renderer.py:153-158 on the GPU: returns the (H, W, 4) Stokes image."""
    own = renderer is None
    rr = renderer or Renderer()
    try:
        db = rr.upload(brick, "f32")
        out = rr.render(scene, db, precision)
        rr.release(db)
        return out["iquv"].astype(np.float64)
    finally:
        if own:
            rr.close()
