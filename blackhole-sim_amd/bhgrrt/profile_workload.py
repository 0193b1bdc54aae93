"""Fixed workload for rocprofv3 runs: renders of the 1024^2 Kerr Stokes scene plus a calibration copy.

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- \
        python3 -m bhgrrt.profile_workload --paths f32:corner:fp32,bf16:corner:mixed
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o run -- python3 -m bhgrrt.profile_workload

Each path is storage:layout:precision; every path is rendered twice (the second
launch is the measured one).  A 256 MiB device-to-device copy follows, so the
FETCH_SIZE / WRITE_SIZE counters can be scaled by what they read for a copy of
known size.
"""

from __future__ import annotations

import argparse
import sys
from pathlib import Path

_ROOT = Path(__file__).resolve().parents[2]
for _p in (_ROOT, _ROOT / "blackhole-sim_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

from bhgrrt import _lib, scenes  # noqa: E402
from bhgrrt.render import Renderer  # noqa: E402

PREC = {"fp64": 0, "fp32": 1, "mixed": 2}
COPY_BYTES = 256 << 20


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--paths", default="f32:corner:fp32,bf16:corner:mixed,f16:corner:mixed,f32:soa:fp32")
    ap.add_argument("--size", type=int, default=1024)
    args = ap.parse_args(argv)
    sc, br = scenes.scale_plain_scene(args.size, args.size), scenes.scale_brick()
    r = Renderer(0)
    n = sc.rays
    bufs = (r.alloc(n * 32), r.alloc(n * 4), r.alloc(n))
    for spec in args.paths.split(","):
        storage, layout, prec = spec.split(":")
        db = r.upload(br, storage, layout)
        for _ in range(2):
            r.render_device(sc, db, PREC[prec], *bufs)
            r.synchronize()
        steps, samples = r.last_work()
        print(f"{spec}: {r.last_kernel_ms():.3f} ms, {steps} RK2 steps, {samples} brick samples", flush=True)
        r.release(db)
    a, b = r.alloc(COPY_BYTES), r.alloc(COPY_BYTES)
    _lib.check(_lib.lib().bh_copy_device(r.ctx, b.ptr, a.ptr, COPY_BYTES))
    print(f"calibration copy: {COPY_BYTES} bytes", flush=True)
    r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
