"""bhgrrt: MI355X-native Kerr Stokes GRRT hot path (RK2 geodesics, brick sampler, IQUV transfer).

Host code in Python over a ctypes C ABI (include/bhgrrt.h) to hand-written HIP
kernels for gfx950.  Submodules:

* ``scenes``         -- camera / scene / brick descriptions and seeded generators
* ``native_kernels`` -- mirror of the reference's native_kernels.py API (FP64 micro-kernels)
* ``render``         -- fused render API (FP64, FP32, mixed precision), batches
* ``accelerated``    -- mirror of accelerated_renderer.py: config, image and progressive renders
* ``surrogate``      -- learned surrogate (MFMA MLP, training, knot render)
* ``autotune``       -- launch-shape autotuner for the persistent kernels
* ``grmhd``          -- GRMHD snapshot fields and the analytic torus fixture
* ``coefficient_bricks`` -- coefficient-brick precompute (synchrotron model on the GPU)
"""

__version__ = "0.9.0"
