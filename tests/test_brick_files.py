"""Brick files in the reference's npz schema (coefficient_bricks.py:57-75): written by Brick.save_npz or
laid out as the reference writes them (float16 / float64 coefficients included), read back by
Brick.load_npz without unpickling."""

import numpy as np
import pytest

from bhgrrt import scenes

KEYS = ("r", "theta", "phi", "coeffs", "coeff_names", "nu_hz", "spin_a", "dtype_name")


def test_save_load_round_trip(tmp_path):
    br = scenes.make_brick(6, 5, 8, spin=0.5, seed=3)
    br.meta.update(nu_hz=345e9, spin_a=0.5)
    p = tmp_path / "b.npz"
    br.save_npz(p)
    with np.load(p, allow_pickle=False) as z:
        assert set(z.files) == set(KEYS)
        assert list(z["coeff_names"]) == list(scenes.COEFF_NAMES)
        assert str(z["dtype_name"]) == "float32"
    back = scenes.Brick.load_npz(p)
    for k in ("r", "theta", "phi", "coeffs"):
        np.testing.assert_array_equal(getattr(back, k), getattr(br, k))
    assert back.meta == {"nu_hz": 345e9, "spin_a": 0.5}


@pytest.mark.parametrize("dtype", ["float16", "float64"])
def test_reference_layout_with_other_coefficient_dtypes(tmp_path, dtype):
    """precompute_coefficient_bricks may store float16 or float64 coefficients (coefficient_bricks.py:119-152);
    they load as float32 (exact for float16, rounded to nearest for float64)."""
    src = scenes.make_brick(5, 4, 6, spin=0.9, seed=9)
    coeffs = src.coeffs.astype(dtype)
    p = tmp_path / "ref.npz"
    np.savez_compressed(p, r=src.r, theta=src.theta, phi=src.phi, coeffs=coeffs,
                        coeff_names=np.array(scenes.COEFF_NAMES), nu_hz=np.array(230e9), spin_a=np.array(0.9),
                        dtype_name=np.array(dtype))
    back = scenes.Brick.load_npz(p)
    assert back.coeffs.dtype == np.float32
    np.testing.assert_array_equal(back.coeffs, coeffs.astype(np.float32))
    assert back.meta["spin_a"] == 0.9


def test_bad_files_are_rejected(tmp_path):
    src = scenes.make_brick(5, 4, 6, spin=0.9, seed=9)
    p = tmp_path / "bad.npz"
    np.savez_compressed(p, r=src.r, theta=src.theta, phi=src.phi, coeffs=src.coeffs[:, :, :, :10],
                        coeff_names=np.array(scenes.COEFF_NAMES[:10]), nu_hz=np.array(230e9), spin_a=np.array(0.9),
                        dtype_name=np.array("float32"))
    with pytest.raises(ValueError, match="coeffs must have shape"):
        scenes.Brick.load_npz(p)
    c = src.coeffs.copy()
    c[0, 0, 0, 0] = np.nan
    np.savez_compressed(p, r=src.r, theta=src.theta, phi=src.phi, coeffs=c, coeff_names=np.array(scenes.COEFF_NAMES),
                        nu_hz=np.array(230e9), spin_a=np.array(0.9), dtype_name=np.array("float32"))
    with pytest.raises(ValueError, match="non-finite"):
        scenes.Brick.load_npz(p)
