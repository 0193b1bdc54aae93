"""The diagnostics the reference's doctor and benchmark read (platform_probe.py:68-97,
native_loader.py:31-61, benchmark.py:32-45).

On the CPU: the drop-in ``blackhole_native`` answers ``native_core_status``'s two probes
(``core_version()`` = "0.9.0", lib.rs:19-22; ``detect_arch()`` with string arch/os/family/pointer_width,
lib.rs:24-35) and ``bhgrrt.doctor.doctor_report`` reports the core.  On the GPU: the device is listed
and the benchmark rows have the reference's ``BenchmarkResult`` fields.
"""

import platform

import pytest

import blackhole_native
from bhgrrt import doctor


def test_native_core_probes_match_the_reference_contract():
    assert blackhole_native.core_version() == "0.9.0"
    arch = blackhole_native.detect_arch()
    assert set(arch) == {"arch", "os", "family", "pointer_width"} and all(isinstance(v, str) for v in arch.values())
    assert arch["arch"] == platform.machine().lower().replace("amd64", "x86_64")
    assert arch["pointer_width"] == "64"


def test_doctor_report_on_any_host():
    rep = doctor.doctor_report()
    assert rep["native_core_loaded"] is True and rep["native_core_version"] == "0.9.0"
    assert rep["native_core_arch"] == blackhole_native.detect_arch()["arch"]
    assert isinstance(rep["devices"], list) and isinstance(rep["warnings"], tuple)
    if not rep["devices"]:
        assert any("no HIP device" in w for w in rep["warnings"])


def test_benchmark_rows_reject_bad_sizes():
    with pytest.raises(ValueError):
        doctor.gpu_benchmark_rows(nr=0)


@pytest.mark.gpu
def test_doctor_sees_the_gpu_and_benchmark_rows_have_the_reference_fields():
    rep = doctor.doctor_report()
    assert rep["gpu_backend"] == "rocm" and rep["devices"] and rep["devices"][0]["compute_units"] > 0
    print(rep["devices"])
    rows = doctor.gpu_benchmark_rows(nr=3, ntheta=3, nphi=4, point_count=7, iterations=2)
    assert [r["name"] for r in rows] == ["coefficient_brick_precompute", "sample_brick_trilinear_native",
                                         "stokes_rk2_brick_native"]
    assert [r["items"] for r in rows] == [72, 14, 72]
    for r in rows:
        assert set(r) == {"name", "seconds", "iterations", "items", "items_per_second", "metadata"}
        assert r["seconds"] > 0 and r["items_per_second"] > 0 and r["metadata"]["backend"] == "hip-gfx950"
