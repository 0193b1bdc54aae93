"""CPU FP64 oracle for the Kerr Stokes hot path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference's algorithms in NumPy so the HIP path can
be checked against them.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline; the product path never routes through it.

Modules and the reference code each follows:

* ``oracle.micro``  -- sampler brackets, trilinear sampling, valid mask, Stokes
  RK2 step and the invalid-sample policy (python/blackhole_sim/native_kernels.py,
  native/core/src/lib.rs:37-366).
* ``oracle.kerr``   -- Kerr metric, analytic metric derivatives, Hamiltonian
  RHS and the ZAMO camera (python/blackhole_sim/kerr.py:27-313).
* ``oracle.render`` -- the fused per-ray render: RK2 midpoint geodesic steps
  (scheme of native/rocm/kerr_stokes_kernel.hip:13, termination of
  kerr.py:357-377) composed with the brick transfer of
  python/blackhole_sim/accelerated_renderer.py:100-126.

Pinning: every function here is checked in tests/test_oracle_*.py against the
reference's own known-answer tests (lib.rs:407-464,
test_native_sampler_parity.py, test_native_stokes_parity.py, test_kerr.py,
test_geodesics.py, test_synchrotron_polarized_transfer.py) and against closed
forms and invariants (metric inverse identity, null condition, finite-difference
derivatives, Carter constant, RK2 order, constant-coefficient transfer, Bardeen
critical curve).
"""
