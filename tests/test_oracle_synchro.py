"""The coefficient-brick oracle against the reference's own outputs and the closed forms it must obey.

tests/golden/synchrotron_vectors.npz holds outputs of the reference's precompute_coefficient_bricks
(coefficient_bricks.py:118-152) on the analytic tori of its tests and benchmark, its hybrid model
(synchrotron.py:315-325) on 32 plasma frames at 230 and 86 GHz, and its kernel table
(synchrotron.py:158-163); make_synchrotron_vectors.py beside it wrote them.
"""

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import synchro as osy

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "synchrotron_vectors.npz")


def _close(a, b, rtol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = np.maximum(np.abs(b), 1e-300)
    bad = np.abs(a - b) > rtol * scale
    assert not bad.any(), f"{bad.sum()} of {bad.size} differ; worst {np.max(np.abs(a - b) / scale):.2e}"


def test_kernel_table_matches_reference():
    xs, F, G = osy.kernel_table()
    np.testing.assert_array_equal(xs, GOLD["table_x"])
    _close(F, GOLD["table_F"], 1e-15)
    _close(G, GOLD["table_G"], 1e-15)


@pytest.mark.parametrize("k", [0, 1])
def test_hybrid_model_matches_reference_frames(k):
    fr, nu = GOLD["frames"], float(GOLD["frame_nu"][k])
    got = osy.hybrid_coefficients(fr[:, 0], fr[:, 1], fr[:, 2], fr[:, 3], nu)
    ref = GOLD["frame_coeffs"][k]
    # zero entries (j_v, alpha_v, rho_v; empty cells) must be zero, the rest agree to rounding
    np.testing.assert_array_equal(got == 0.0, ref == 0.0)
    _close(got[ref != 0.0], ref[ref != 0.0], 1e-12)


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_brick_precompute_matches_reference(tag):
    sc = GOLD[f"{tag}_scaling"]
    stride = int(GOLD[f"{tag}_stride"])
    for key in [k for k in GOLD.files if k.startswith(f"{tag}_coeffs_")]:
        dt = key.rsplit("_", 1)[1]
        got = osy.precompute_coefficient_bricks(GOLD[f"{tag}_rho"], GOLD[f"{tag}_theta_e"], GOLD[f"{tag}_b_con"],
                                                sc[0], sc[1], 230e9, dt, stride, sc[2])
        ref = GOLD[key]
        assert got.dtype == ref.dtype and got.shape == ref.shape
        if dt == "float64":
            np.testing.assert_array_equal(got == 0.0, ref == 0.0)
            _close(got[ref != 0.0], ref[ref != 0.0], 1e-12)
        else:
            # the FP64 values agree to ~1e-15, so the narrow casts agree except at a rounding tie
            assert np.mean(got == ref) >= 0.999
            _close(got.astype(float)[ref != 0], ref.astype(float)[ref != 0], 2.0 ** -10)


def _frames(n=40, seed=5):
    rng = np.random.default_rng(seed)
    return (10 ** rng.uniform(2, 8, n), 10 ** rng.uniform(-1.5, 2, n), 10 ** rng.uniform(-1, 3, n),
            rng.uniform(-3, 3, n))


def test_thermal_part_obeys_kirchhoff():
    ne, th, b, ev = _frames()
    nu = 230e9
    c = osy.thermal_coefficients(ne, th, b, ev, nu)
    bnu = osy.planck_nu(nu, th)
    # alpha = j / B_nu for I and for the rotated Q/U pair
    _close(c[:, 4] * bnu, c[:, 0], 1e-14)
    ok = c[:, 1] != 0
    _close(c[ok, 5] * bnu[ok], c[ok, 1], 1e-14)


def test_evpa_rotation_invariants():
    ne, th, b, ev = _frames()
    nu = 86e9
    c0 = osy.hybrid_coefficients(ne, th, b, np.zeros_like(ev), nu)
    c = osy.hybrid_coefficients(ne, th, b, ev, nu)
    # chi = 0 leaves no U terms; rotation keeps I, V, alpha_I, rho_V and the Q/U lengths, and turns by 2 chi
    assert np.all(c0[:, [2, 6, 10]] == 0)
    for k in (0, 3, 4, 7, 8):
        np.testing.assert_array_equal(c[:, k], c0[:, k])
    for q, u in ((1, 2), (5, 6), (9, 10)):
        _close(np.hypot(c[:, q], c[:, u]), np.abs(c0[:, q]), 1e-14)
        ang = np.arctan2(c[:, u], c[:, q])
        want = np.angle(np.exp(1j * (2 * ev + np.where(c0[:, q] < 0, math.pi, 0.0))))
        assert np.max(np.abs(np.angle(np.exp(1j * (ang - want))))) < 1e-12


def test_coefficients_are_linear_in_density():
    ne, th, b, ev = _frames()
    c1 = osy.hybrid_coefficients(ne, th, b, ev, 230e9)
    c2 = osy.hybrid_coefficients(2.0 * ne, th, b, ev, 230e9)
    np.testing.assert_array_equal(c2, 2.0 * c1)


def test_no_plasma_no_coefficients():
    c = osy.hybrid_coefficients(np.array([0.0, 1e6, 1e6]), np.array([5.0, 5.0, 5.0]), np.array([10.0, 0.0, 10.0]),
                                np.zeros(3), 230e9)
    assert np.all(c[:2] == 0) and np.all(c[2, [0, 4, 9]] > 0)
