"""bh_render_device_noalloc (counters and tile schedule in the library's static per-GPU array) against
bh_render_device, through the C ABI, on two calls in a row with the second image smaller.

Both entry points run the same kernels on the same schedule, and a lone ray split over two waves rounds
every step as the unsplit ray does, so statuses, step counts, the work counters and IQUV must be equal.
"""

import numpy as np
import pytest

from bhgrrt import scenes
from bhgrrt.render import PREC_FP32, PREC_FP64, PREC_MIXED

pytestmark = pytest.mark.gpu


def _run(renderer, fn, sc, db, precision):
    n = sc.rays
    esz = 8 if precision == PREC_FP64 else 4
    bufs = [renderer.alloc(n * 4 * esz), renderer.alloc(n * 4), renderer.alloc(n)]
    try:
        fn(sc, db, precision, *bufs)
        renderer.synchronize()
        work = renderer.last_work()
        iquv = bufs[0].to_host(np.empty((sc.height, sc.width, 4), np.float64 if esz == 8 else np.float32))
        steps = bufs[1].to_host(np.empty((sc.height, sc.width), np.int32))
        status = bufs[2].to_host(np.empty((sc.height, sc.width), np.uint8))
    finally:
        for b in bufs:
            b.free()
    return iquv, steps, status, work


@pytest.mark.parametrize("storage,layout,precision", [("f32", "corner", PREC_FP32), ("bf16", "corner", PREC_MIXED),
                                                      ("f32", "soa", PREC_FP64)])
def test_noalloc_render_matches_render_device(renderer, storage, layout, precision):
    br = scenes.scale_brick()
    db = renderer.upload(br, storage, layout)
    try:
        for w, h in ((384, 320), (136, 72)):   # the second call is smaller than the first
            sc = scenes.scale_plain_scene(w, h)
            print(f"noalloc {storage}/{layout} prec {precision} {w}x{h}", flush=True)
            ref = _run(renderer, renderer.render_device, sc, db, precision)
            got = _run(renderer, renderer.render_device_noalloc, sc, db, precision)
            np.testing.assert_array_equal(got[2], ref[2])
            np.testing.assert_array_equal(got[1], ref[1])
            assert got[3] == ref[3]
            np.testing.assert_array_equal(got[0], ref[0])
    finally:
        renderer.release(db)


def test_noalloc_render_rejects_images_beyond_its_scratch(renderer):
    br = scenes.small_brick()
    db = renderer.upload(br, "f32", "corner")
    try:
        sc = scenes.small_plain_scene().with_size(2056, 2048)   # 257 x 256 tiles
        # outputs of the full size, so that nothing could be written out of bounds even if the check failed
        bufs = [renderer.alloc(sc.rays * 16), renderer.alloc(sc.rays * 4), renderer.alloc(sc.rays)]
        with pytest.raises(ValueError, match="no-allocation render holds at most 65536"):
            renderer.render_device_noalloc(sc, db, PREC_FP32, *bufs)
        renderer.synchronize()
        for b in bufs:
            b.free()
    finally:
        renderer.release(db)
