"""Drop-in for the reference's coefficient-brick renderer API (python/blackhole_sim/accelerated_renderer.py).

Same names and argument meaning as the reference: ``AcceleratedRenderConfig`` (:36-46, with the
``PolarizedTransferConfig`` fields the renderer reads, polarized_transfer.py:23-29),
``render_tile_stokes_bricks`` (:129-150), ``render_stokes_image_bricks`` (:153-158),
``render_progressive_stokes_bricks`` (:161-166) and
``ProgressiveFrame`` (:49-54), plus ``progressive_levels`` and ``generate_tiles``
(tiles.py:30-50).  The image comes from the fused HIP render instead of the per-pixel CPU loop:

* one Kerr ray per pixel with the same camera and pixel centres (ndc = 2(i+1/2)/W - 1,
  1 - 2(j+1/2)/H) and the same transfer rule (midpoint sample, ds = L max(|d(r,theta,phi)|, 1e-14),
  RK2 or the exact matrix step, tau cap, nan_to_num);
* the geodesic integrator is the reference tracer's classical RK4 (kerr.py:306-313) with its fixed
  step ``cfg.step`` and termination rules (kerr.py:357-377); ``integrator="rk2"`` selects the RK2
  midpoint scheme of the GPU asset and ``step_mode=1`` the adaptive step h = step (r - r_+);
* the FP64 render samples the float32 brick in float32 arithmetic exactly as the reference's
  sample_coefficient_brick (:73-97) runs under NumPy, so its images equal the reference renderer's
  to rounding of the geodesic arithmetic (statuses and step counts exactly; Stokes within 1e-10 of
  |I| on the reference's own outputs, tests/golden/accelerated_vectors.npz); FP32 and mixed renders
  trace and sample in FP32 (errors stated in tests/test_gpu_accelerated_api.py);
* tiles (``tile_size``) only schedule the reference's CPU loop; the GPU renders the whole image in one
  launch, and ``render_progressive_stokes_bricks`` renders every progressive level in one batched
  launch.

``snapshot`` may be anything with a ``spin_a`` attribute (the reference passes a GRMHDSnapshot) or the
spin as a number.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .render import PREC_FP32, PREC_FP64, PREC_MIXED, Batch, Renderer
from .scenes import STEP_FIXED, Brick, Camera, Scene


@dataclass(frozen=True)
class PolarizedTransferConfig:
    """The transfer settings the brick renderer reads (polarized_transfer.py:23-29)."""

    affine_to_length_cm: float | None = None
    max_optical_depth: float = 30.0
    use_exact_matrix_step: bool = True


@dataclass(frozen=True)
class AcceleratedRenderConfig:
    """accelerated_renderer.py:36-46; ``step_mode`` (0 fixed, 1 adaptive h = step (r - r_+)) and ``integrator``
    ("rk4" as the reference tracer, or "rk2" midpoint) are additions."""

    width: int = 160
    height: int = 90
    camera: Camera = field(default_factory=lambda: Camera.from_degrees(r=55.0, inclination_degrees=65.0,
                                                                       fov_y_degrees=32.0))
    step: float = 0.06
    max_steps: int = 5000
    escape_radius: float = 220.0
    transfer: PolarizedTransferConfig = field(default_factory=lambda: PolarizedTransferConfig(use_exact_matrix_step=False))
    tile_size: int = 64
    progressive_min_width: int = 480
    step_mode: int = STEP_FIXED
    integrator: str = "rk4"


@dataclass(frozen=True)
class ProgressiveFrame:
    """accelerated_renderer.py:49-54."""

    level: int
    width: int
    height: int
    stokes: np.ndarray


@dataclass(frozen=True, order=True)
class RenderTile:
    """tiles.py:9-30."""

    level: int
    y0: int
    x0: int
    x1: int
    y1: int

    @property
    def width(self) -> int:
        return self.x1 - self.x0

    @property
    def height(self) -> int:
        return self.y1 - self.y0


def generate_tiles(width: int, height: int, tile_size: int = 64, level: int = 0) -> list[RenderTile]:
    """tiles.py:30-35: row-major tiles of ``tile_size``, clipped at the image edge."""
    return [RenderTile(level, x0=x, y0=y, x1=min(x + tile_size, width), y1=min(y + tile_size, height))
            for y in range(0, int(height), int(tile_size)) for x in range(0, int(width), int(tile_size))]


def progressive_levels(width: int, height: int, min_width: int = 480) -> list[tuple[int, int, int]]:
    """tiles.py:38-50: (level, width, height) from coarse to full; each level doubles the resolution and
    the coarsest is the smallest power-of-two reduction whose width stays >= min_width."""
    w = int(width)
    scale = 1
    while w // (scale * 2) >= min_width:
        scale *= 2
    out = []
    level = 0
    while scale >= 1:
        out.append((level, max(1, int(width) // scale), max(1, int(height) // scale)))
        level += 1
        scale //= 2
    return out


def _spin(snapshot) -> float:
    return float(getattr(snapshot, "spin_a", snapshot))


def scene_from_config(cfg: AcceleratedRenderConfig, snapshot, width: int | None = None,
                      height: int | None = None) -> Scene:
    """The Scene this build renders for ``cfg`` (at ``width`` x ``height`` when given)."""
    t = cfg.transfer
    return Scene(spin=_spin(snapshot), camera=cfg.camera, width=int(width or cfg.width), height=int(height or cfg.height),
                 step_mode=int(cfg.step_mode), step=float(cfg.step), max_steps=int(cfg.max_steps),
                 escape_radius=float(cfg.escape_radius), affine_to_length_cm=float(t.affine_to_length_cm or 1.0),
                 max_optical_depth=float(t.max_optical_depth), exact_step=bool(t.use_exact_matrix_step),
                 integrator=str(cfg.integrator), numpy32_sampling=True)


def _upload(rr: Renderer, bricks: Brick, precision: int):
    if precision == PREC_FP64:
        return rr.upload(bricks, "f32", "soa")
    return rr.upload(bricks, "bf16" if precision == PREC_MIXED else "f32", "corner")


def render_stokes_image_bricks(cfg: AcceleratedRenderConfig, snapshot, bricks: Brick, precision: int = PREC_FP64,
                               renderer: Renderer | None = None) -> np.ndarray:
    """accelerated_renderer.py:153-158 on the GPU: the (H, W, 4) float64 Stokes image."""
    own = renderer is None
    rr = renderer or Renderer()
    try:
        db = _upload(rr, bricks, precision)
        try:
            out = rr.render(scene_from_config(cfg, snapshot), db, precision)
        finally:
            rr.release(db)
        return out["iquv"].astype(np.float64)
    finally:
        if own:
            rr.close()


def render_tile_stokes_bricks(cfg: AcceleratedRenderConfig, snapshot, bricks: Brick, tile: RenderTile,
                              precision: int = PREC_FP64, renderer: Renderer | None = None) -> np.ndarray:
    """accelerated_renderer.py:129-150: the (tile.height, tile.width, 4) block of the image.  The GPU traces
    the whole image in one launch and the block is cut from it, so callers that loop over tiles should
    call ``render_stokes_image_bricks`` once instead."""
    img = render_stokes_image_bricks(cfg, snapshot, bricks, precision, renderer)
    return img[tile.y0:tile.y1, tile.x0:tile.x1].copy()


def render_progressive_stokes_bricks(cfg: AcceleratedRenderConfig, snapshot, bricks: Brick,
                                     precision: int = PREC_FP64, renderer: Renderer | None = None
                                     ) -> list[ProgressiveFrame]:
    """accelerated_renderer.py:161-166 on the GPU: every level of ``progressive_levels`` traced by one
    batched launch; frames coarse to full, each (h, w, 4) float64."""
    own = renderer is None
    rr = renderer or Renderer()
    try:
        levels = progressive_levels(cfg.width, cfg.height, min_width=cfg.progressive_min_width)
        db = _upload(rr, bricks, precision)
        batch = Batch(rr, [(scene_from_config(cfg, snapshot, w, h), db) for _, w, h in levels], precision)
        try:
            batch.launch(graph=False)
            results = batch.results()
        finally:
            batch.close()
            rr.release(db)
        return [ProgressiveFrame(level=lv, width=w, height=h, stokes=res["iquv"].astype(np.float64))
                for (lv, w, h), res in zip(levels, results)]
    finally:
        if own:
            rr.close()


__all__ = ["AcceleratedRenderConfig", "PolarizedTransferConfig", "ProgressiveFrame", "RenderTile", "generate_tiles",
           "progressive_levels", "scene_from_config", "render_stokes_image_bricks", "render_tile_stokes_bricks",
           "render_progressive_stokes_bricks",
           "PREC_FP32", "PREC_FP64", "PREC_MIXED"]
