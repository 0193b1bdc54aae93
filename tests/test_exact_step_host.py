"""The exact Stokes step of the render kernels (csrc/bh_stokes_exact.h), built for the host, on the CPU.

The same header is compiled into the GPU kernels; here it runs on the CPU so that the inputs a render
must never hang on can be tried without a GPU: ds = +inf, NaN and 1e308, and absorption of 1e300.
Semantics follow polarized_transfer.py:51-70 and the oracle (scipy's expm, the reference's routine):
S unchanged for ds <= 0; NaN for a non-finite alpha ds (scipy's expm returns NaN there, and a NaN ds
passes the reference's ds <= 0 test); otherwise the exact step, within 1e-12 of the oracle in FP64.
A huge but finite alpha ds ends after at most ~1100 halvings and gives the step's limit j / K.
"""

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import micro

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc is needed to build the host copy of the exact step")
    out = tmp_path_factory.mktemp("exact") / "libexacthost.so"
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-fPIC", "-shared", "-std=c++17",
                    f"-I{ROOT / 'blackhole-sim_amd' / 'csrc'}", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "native" / "exact_step_host.hip"), "-o", str(out)], check=True,
                   capture_output=True)
    L = C.CDLL(str(out))
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    L.bh_test_exact_halvings.argtypes, L.bh_test_exact_halvings.restype = [dp, C.c_double], C.c_int
    L.bh_test_stokes_exact.argtypes, L.bh_test_stokes_exact.restype = [dp, C.c_double, dp], None
    L.bh_test_stokes_exact_f32.argtypes, L.bh_test_stokes_exact_f32.restype = [fp, C.c_float, fp], None
    return L


def _step(lib, c, ds, s):
    c = np.ascontiguousarray(c, dtype=np.float64)
    out = np.array(s, dtype=np.float64)
    lib.bh_test_stokes_exact(c.ctypes.data_as(C.POINTER(C.c_double)), float(ds), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def _halvings(lib, c, ds):
    c = np.ascontiguousarray(c, dtype=np.float64)
    return lib.bh_test_exact_halvings(c.ctypes.data_as(C.POINTER(C.c_double)), float(ds))


def _coeffs(rng, scale=1.0):
    a_i = rng.uniform(0.1, 2.0) * scale
    c = np.empty(11)
    c[0] = rng.uniform(0.1, 2.0)
    c[1:4] = rng.uniform(-0.5, 0.5, 3) * c[0]
    c[4] = a_i
    c[5:8] = rng.uniform(-0.4, 0.4, 3) * a_i
    c[8:11] = rng.normal(0.0, 0.5, 3) * a_i
    return c


def test_exact_step_matches_oracle(lib):
    rng = np.random.default_rng(7)
    for _ in range(200):
        c, s = _coeffs(rng), rng.normal(size=4)
        ds = float(np.exp(rng.uniform(np.log(1e-3), np.log(60.0))))
        got = _step(lib, c, ds, s)
        ref = micro.stokes_exact_step(s[None], c[None], np.array([ds]))[0]
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # FP32 build of the same code (the FP32 kernels' rule, TAYLOR 8)
    c, s, ds = _coeffs(rng), rng.normal(size=4), 0.7
    c32, s32 = c.astype(np.float32), s.astype(np.float32)
    lib.bh_test_stokes_exact_f32(c32.ctypes.data_as(C.POINTER(C.c_float)), ds, s32.ctypes.data_as(C.POINTER(C.c_float)))
    ref = micro.stokes_exact_step(s[None], c[None], np.array([ds]))[0]
    np.testing.assert_allclose(s32, ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max())


@pytest.mark.parametrize("ds", [0.0, -1.0, -np.inf, -1e308])
def test_non_positive_ds_leaves_the_stokes_vector(lib, ds):
    c, s = _coeffs(np.random.default_rng(1)), np.array([0.3, -0.1, 0.2, 0.05])
    np.testing.assert_array_equal(_step(lib, c, ds, s), s)
    np.testing.assert_array_equal(micro.stokes_exact_step(s[None], c[None], np.array([ds]))[0], s)


@pytest.mark.parametrize("ds", [np.inf, np.nan])
def test_non_finite_ds_gives_nan_at_once(lib, ds):
    c, s = _coeffs(np.random.default_rng(2)), np.array([0.3, -0.1, 0.2, 0.05])
    assert _halvings(lib, c, ds) == -1
    assert np.all(np.isnan(_step(lib, c, ds, s)))
    assert np.all(np.isnan(micro.stokes_exact_step(s[None], c[None], np.array([ds]))[0]))


def test_non_finite_coefficient_gives_nan(lib):
    c = _coeffs(np.random.default_rng(3))
    c[6] = np.inf
    assert _halvings(lib, c, 0.5) == -1
    assert np.all(np.isnan(_step(lib, c, 0.5, np.zeros(4))))


def test_overflowing_alpha_ds_ends_at_once(lib):
    """alpha ds beyond the double range (ds = 1e308, alpha ~ 1): the norm is infinite, NaN as scipy gives."""
    c, s = _coeffs(np.random.default_rng(4)), np.zeros(4)
    assert _halvings(lib, c, 1e308) == -1
    assert np.all(np.isnan(_step(lib, c, 1e308, s)))
    assert np.all(np.isnan(micro.stokes_exact_step(s[None], c[None], np.array([1e308]))[0]))


@pytest.mark.parametrize("scale,ds", [(1e300, 1.0), (1e-300, 1e308), (1.0, 1e300)])
def test_huge_finite_alpha_ds_ends_with_the_limit(lib, scale, ds):
    """A finite alpha ds up to ~1e308: at most ~1030 halvings, then squaring; the result is the step's
    limit for alpha ds -> inf, S = K^-1 j (the emission-absorption balance), whatever S was."""
    c = _coeffs(np.random.default_rng(5), scale)
    c[8:11] = 0.0   # no Faraday terms: K = alpha_I I + the symmetric dichroism block, invertible
    K = micro.propagation_matrix(c[None])[0]
    limit = np.linalg.solve(K, c[:4])
    p = _halvings(lib, c, ds)
    assert 0 < p <= 1100
    got = _step(lib, c, ds, np.array([5.0, 1.0, -2.0, 0.5]))
    np.testing.assert_allclose(got, limit, rtol=1e-9, atol=1e-9 * np.abs(limit).max())
