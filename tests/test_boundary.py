"""The C ABI library loads, exports every symbol include/bhgrrt.h declares, and
validates inputs on the host with the reference's ValueError messages (no GPU
compute is issued by these calls)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from bhgrrt import _lib
import blackhole_native
from bhgrrt import native_kernels as nk

HEADER = Path(__file__).resolve().parents[1] / "include" / "bhgrrt.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bh_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(s for s, _, _ in _lib.SIGNATURES) == names


def test_module_identity():
    assert blackhole_native.core_version() == "0.9.0"
    arch = blackhole_native.detect_arch()
    assert set(arch) == {"arch", "os", "family", "pointer_width"}
    assert arch["pointer_width"] == "64"


def test_stokes_validation_messages():
    # lib.rs:228-256
    with pytest.raises(ValueError, match="ds_cm must be finite and non-negative"):
        blackhole_native.stokes_rk2_brick([0.0] * 11, -1.0, [0.0] * 4)
    with pytest.raises(ValueError, match="one or more flat 11-coefficient cells"):
        blackhole_native.stokes_rk2_brick([0.0] * 10, 0.1, [0.0] * 4)
    with pytest.raises(ValueError, match="only finite"):
        blackhole_native.stokes_rk2_brick([float("nan")] + [0.0] * 10, 0.1, [0.0] * 4)
    with pytest.raises(ValueError, match="initial must contain either 4 values"):
        blackhole_native.stokes_rk2_brick([0.0] * 22, 0.1, [0.0] * 5)


def test_sampler_validation_messages():
    # lib.rs:112-169
    c = [0.0] * (2 * 2 * 2 * 11)
    g = [0.0, 1.0]
    with pytest.raises(ValueError, match="r grid must contain at least 2 values"):
        blackhole_native.sample_brick_trilinear(c, [0.0], g, g, [0.5, 0.5, 0.5])
    with pytest.raises(ValueError, match="theta grid must be strictly increasing"):
        blackhole_native.sample_brick_trilinear(c, g, [1.0, 0.0], g, [0.5, 0.5, 0.5])
    with pytest.raises(ValueError, match="exactly r\\*theta\\*phi\\*11 values; expected 88, got 87"):
        blackhole_native.sample_brick_trilinear(c[:-1], g, g, g, [0.5, 0.5, 0.5])
    with pytest.raises(ValueError, match="flat r, theta, phi triples"):
        blackhole_native.sample_brick_trilinear(c, g, g, g, [0.5, 0.5])
    with pytest.raises(ValueError, match="initial must contain either 4 values or 4 values per sample point"):
        blackhole_native.sample_and_step_stokes(c, g, g, g, [0.5, 0.5, 0.5], 0.1, [0.0] * 3)
    with pytest.raises(ValueError, match="invalid_policy"):
        _lib.check(_lib.lib().bh_sample_and_step_stokes(
            _lib.dptr(np.zeros(88)), 88, _lib.dptr(np.array(g)), 2, _lib.dptr(np.array(g)), 2,
            _lib.dptr(np.array(g)), 2, _lib.dptr(np.full(3, 0.5)), 3, 0.1, _lib.dptr(np.zeros(4)), 4, 7,
            _lib.dptr(np.zeros(4))))


def test_python_wrapper_validation_messages():
    # native_kernels.py:53-130 (messages the reference's tests match)
    coeffs = nk.deterministic_stokes_coefficients(2, 2, 2)
    with pytest.raises(ValueError, match="shape"):
        nk.stokes_rk2_brick(coeffs[..., :10], ds_cm=0.05)
    with pytest.raises(ValueError, match="initial"):
        nk.stokes_rk2_brick(coeffs, ds_cm=0.05, initial=np.zeros((2, 4)))
    g = np.array([0.0, 1.0])
    with pytest.raises(ValueError, match="invalid_policy"):
        nk.sample_and_step_stokes(np.zeros((2, 2, 2, 11)), g, g, g, [[0.5, 0.5, 0.5]], 0.05, invalid_policy="break")


def test_no_gpu_is_a_clear_error_not_a_fallback():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = _lib.lib().bh_ctx_create(0, C.byref(h))
    assert rc == _lib.BH_ERR_DEVICE
    assert "no HIP device" in _lib.last_error()
    with pytest.raises(RuntimeError, match="no HIP device"):
        blackhole_native.stokes_rk2_brick([0.0] * 11, 0.1, [0.0] * 4)
