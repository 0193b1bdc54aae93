"""Per-pixel condition of a render, measured with the FP64 oracle itself.

A ray whose image value moves by a large factor when its initial state moves by a rounding error cannot be
held to a tight bar in any precision: two correct FP64 codes that round differently disagree there.  The
small plain scene (64 fixed RK2 steps of h = 20 M from r = 1000) is such a case for most of its rays, whose
steps jump across the coordinate axis near the hole.  ``well_conditioned`` renders the scene with the
oracle at the camera radius and field of view scaled by 1 - delta and 1 + delta and returns the pixels
whose status and step count stay the same and whose I moves by less than ``tol`` relative: on those, a
correct implementation with rounding errors of order delta is within about tol of the oracle.
"""

from __future__ import annotations

import dataclasses

import numpy as np

from oracle import render as orender


def well_conditioned(scene, brick, delta: float, tol: float, ref=None):
    ref = ref if ref is not None else orender.render_image(scene, brick)
    I0 = ref["iquv"][..., 0]
    ok = np.ones(I0.shape, dtype=bool)
    for sgn in (-1.0, 1.0):
        cam = dataclasses.replace(scene.camera, r=scene.camera.r * (1.0 + sgn * delta),
                                  fov_y_degrees=scene.camera.fov_y_degrees * (1.0 + sgn * delta))
        o = orender.render_image(dataclasses.replace(scene, camera=cam), brick)
        rel = np.abs(o["iquv"][..., 0] - I0) / np.maximum(np.abs(I0), 1e-300)
        ok &= (o["status"] == ref["status"]) & (o["steps"] == ref["steps"]) & (rel < tol)
    return ok
