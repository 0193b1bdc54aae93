"""ctypes binding of libbhgrrt.so (C ABI declared in include/bhgrrt.h).

The library is loaded from this package directory.  There is no CPU fallback:
if the library is missing, ``lib()`` raises ImportError, and on a machine
without a gfx950 GPU every device entry point returns BH_ERR_DEVICE, which
``check`` turns into RuntimeError.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libbhgrrt.so"

BH_OK, BH_ERR_VALUE, BH_ERR_DEVICE, BH_ERR_MEMORY, BH_ERR_STATE = 0, -1, -2, -3, -4
POLICY = {"nan": 0, "initial": 1, "zero": 2}
STORE = {"f32": 0, "bf16": 1, "f16": 2}
LAYOUT = {"soa": 0, "corner": 1}
PREC = {"fp64": 0, "fp32": 1, "mixed": 2}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)


class SceneDesc(C.Structure):
    """bh_scene_desc."""
    _fields_ = [
        ("spin", C.c_double), ("cam_r", C.c_double), ("cam_theta", C.c_double), ("cam_phi", C.c_double),
        ("fov_y_degrees", C.c_double), ("roll_degrees", C.c_double),
        ("width", C.c_int32), ("height", C.c_int32), ("step_mode", C.c_int32), ("max_steps", C.c_int32),
        ("step", C.c_double), ("escape_radius", C.c_double), ("affine_to_length_cm", C.c_double),
        ("max_optical_depth", C.c_double), ("transfer", C.c_int32), ("integrator", C.c_int32),
        ("sampling", C.c_int32), ("reserved", C.c_int32),
    ]


TRANSFER_RK2, TRANSFER_EXACT = 0, 1
INTEGRATOR_RK2, INTEGRATOR_RK4 = 0, 1
SAMPLING_NATIVE, SAMPLING_NUMPY32 = 0, 1


class Tuning(C.Structure):
    """bh_tuning."""
    _fields_ = [("block_threads", C.c_int32), ("blocks_per_cu", C.c_int32), ("steps_per_sync", C.c_int32),
                ("refill_min", C.c_int32), ("queues", C.c_int32), ("order", C.c_int32), ("coop_steps", C.c_int32)]


class BatchItem(C.Structure):
    """bh_batch_item."""
    _fields_ = [("scene", SceneDesc), ("brick_id", C.c_int32), ("reserved", C.c_int32)]


class KnotDesc(C.Structure):
    """bh_knot_desc."""
    _fields_ = [("cam_r", C.c_double), ("step", C.c_double), ("escape_radius", C.c_double), ("r_lo", C.c_double),
                ("r_hi", C.c_double), ("step_mode", C.c_int32), ("max_steps", C.c_int32), ("n_knots", C.c_int32),
                ("reserved", C.c_int32)]


class SurrogateDesc(C.Structure):
    """bh_surrogate_desc."""
    _fields_ = [("in_mul", C.c_float * 4), ("in_add", C.c_float * 4), ("out_mul", _fp), ("out_add", _fp),
                ("n_knots", C.c_int32), ("fallback_precision", C.c_int32), ("crit_distance", C.c_float),
                ("flag_max", C.c_float), ("cap_lo", C.c_float), ("cap_hi", C.c_float)]


class SynchrotronDesc(C.Structure):
    """bh_synchrotron_desc."""
    _fields_ = [("nu_hz", C.c_double), ("rho_cgs_per_code", C.c_double), ("b_gauss_per_code", C.c_double),
                ("mean_molecular_weight_per_electron", C.c_double),
                ("th_gamma_grid_size", C.c_int32), ("th_pitch_nodes", C.c_int32),
                ("th_faraday_rotation_scale", C.c_double), ("th_faraday_conversion_scale", C.c_double),
                ("nt_p", C.c_double), ("nt_gamma_min", C.c_double), ("nt_gamma_max", C.c_double),
                ("nt_fraction", C.c_double), ("nt_gamma_grid_size", C.c_int32), ("nt_pitch_nodes", C.c_int32),
                ("nt_absorption_scale", C.c_double), ("nt_faraday_scale", C.c_double)]


# brick value types of the coefficient precompute (BH_STORE_F32, BH_STORE_F16, BH_STORE_F64)
BRICK_DTYPE = {"float32": 0, "float16": 2, "float64": 3}

# (name, restype, argtypes) for every symbol include/bhgrrt.h declares
SIGNATURES = [
    ("bh_core_version", C.c_char_p, []),
    ("bh_detect_arch", C.c_int, [C.c_char_p, C.c_size_t]),
    ("bh_last_error", C.c_char_p, []),
    ("bh_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("bh_device_name", C.c_int, [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]),
    ("bh_stokes_rk2_brick", C.c_int, [_dp, C.c_size_t, C.c_double, _dp, C.c_size_t, _dp]),
    ("bh_sample_brick_trilinear", C.c_int,
     [_dp, C.c_size_t, _dp, C.c_size_t, _dp, C.c_size_t, _dp, C.c_size_t, _dp, C.c_size_t, _dp]),
    ("bh_sample_and_step_stokes", C.c_int,
     [_dp, C.c_size_t, _dp, C.c_size_t, _dp, C.c_size_t, _dp, C.c_size_t, _dp, C.c_size_t,
      C.c_double, _dp, C.c_size_t, C.c_int, _dp]),
    ("bh_sample_brick_valid_mask", C.c_int, [_dp, C.c_size_t, _dp, C.c_size_t, _dp, C.c_size_t, _u8p]),
    ("bh_ctx_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("bh_ctx_destroy", C.c_int, [C.c_void_p]),
    ("bh_ctx_synchronize", C.c_int, [C.c_void_p]),
    ("bh_ctx_last_kernel_ms", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("bh_device_alloc", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("bh_device_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("bh_copy_to_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("bh_copy_to_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("bh_copy_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("bh_sample_brick_trilinear_device", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
      C.c_void_p, C.c_int64, C.c_void_p]),
    ("bh_brick_upload", C.c_int,
     [C.c_void_p, _fp, _fp, C.c_int32, _fp, C.c_int32, _fp, C.c_int32, C.c_int32, C.c_int32, _i32p]),
    ("bh_brick_release", C.c_int, [C.c_void_p, C.c_int32]),
    ("bh_ctx_set_tuning", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Tuning)]),
    ("bh_ctx_get_tuning", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Tuning)]),
    ("bh_ctx_last_work", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("bh_ctx_last_coop", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("bh_batch_create", C.c_int,
     [C.c_void_p, C.POINTER(BatchItem), C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_void_p)]),
    ("bh_batch_launch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("bh_batch_launch_graph", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("bh_batch_destroy", C.c_int, [C.c_void_p]),
    ("bh_render", C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.c_int32, C.c_int32, C.c_void_p, _i32p, _u8p]),
    ("bh_render_device", C.c_int,
     [C.c_void_p, C.POINTER(SceneDesc), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("bh_render_device_noalloc", C.c_int,
     [C.c_void_p, C.POINTER(SceneDesc), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("bh_mlp_create", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("bh_mlp_destroy", C.c_int, [C.c_void_p]),
    ("bh_mlp_shape", C.c_int, [C.c_void_p, _i32p, _i32p, C.POINTER(C.c_int64)]),
    ("bh_mlp_set_params", C.c_int, [C.c_void_p, _fp]),
    ("bh_mlp_get_params", C.c_int, [C.c_void_p, _fp]),
    ("bh_mlp_get_adam", C.c_int, [C.c_void_p, _fp, _fp, C.POINTER(C.c_int64)]),
    ("bh_mlp_set_adam", C.c_int, [C.c_void_p, _fp, _fp, C.c_int64]),
    ("bh_mlp_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32]),
    ("bh_mlp_set_fused", C.c_int, [C.c_void_p, C.c_int32]),
    ("bh_mlp_gradient", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, _fp, _fp]),
    ("bh_mlp_train_step", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_float, _fp]),
    ("bh_columns_affine", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, _fp, _fp]),
    ("bh_geodesic_knots", C.c_int,
     [C.c_void_p, C.POINTER(KnotDesc), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("bh_surrogate_render_device", C.c_int,
     [C.c_void_p, C.POINTER(SceneDesc), C.c_int32, C.c_void_p, C.POINTER(SurrogateDesc), C.c_void_p, C.c_size_t,
      C.c_void_p, C.c_void_p, C.c_void_p]),
    ("bh_knot_render_device", C.c_int,
     [C.c_void_p, C.POINTER(SceneDesc), C.c_int32, C.c_void_p, C.c_int64, C.POINTER(SurrogateDesc), C.c_void_p,
      C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("bh_surrogate_last_fallback", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("bh_surrogate_work_bytes", C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_size_t)]),
    ("bh_surrogate_batch_create", C.c_int, [C.c_void_p, C.POINTER(BatchItem), C.c_int32, C.c_void_p,
                                            C.POINTER(SurrogateDesc), C.POINTER(C.c_int64), C.POINTER(C.c_void_p)]),
    ("bh_surrogate_batch_launch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    ("bh_surrogate_batch_fallback", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("bh_surrogate_batch_destroy", C.c_int, [C.c_void_p]),
    ("bh_precompute_coefficient_bricks", C.c_int,
     [_dp, _dp, _dp, C.c_int64, C.POINTER(SynchrotronDesc), C.c_int32, C.c_void_p]),
    ("bh_precompute_coefficient_bricks_device", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(SynchrotronDesc), C.c_int32,
      C.c_void_p]),
    ("bh_bessel_k", C.c_int, [C.c_double, C.c_double, _dp]),
    ("bh_gauss_legendre", C.c_int, [C.c_int32, _dp, _dp]),
]

_lock = threading.Lock()
_lib = None


def lib() -> C.CDLL:
    """Load libbhgrrt.so once; raises ImportError when it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                                  f"g.build()'` (hipcc, gfx950); there is no CPU fallback")
            handle = C.CDLL(str(LIB_PATH))
            for name, res, args in SIGNATURES:
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().bh_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Raise ValueError (reference semantics) or RuntimeError for a non-zero status."""
    if rc == BH_OK:
        return
    msg = last_error() or f"bhgrrt error {rc}"
    if rc == BH_ERR_VALUE:
        raise ValueError(msg)
    raise RuntimeError(msg)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def device_count() -> int:
    n = C.c_int(0)
    lib().bh_device_count(C.byref(n))
    return int(n.value)
