"""Host-side pieces of the coefficient precompute, without a GPU.

* ``bh_bessel_k`` (trapezoidal rule on the cosh integral) against scipy.special.kv, which the
  reference calls for the x K_{2/3}(x) table (synchrotron.py:148) and the Maxwell-Juttner
  normalisation K_2 (synchrotron.py:195), over the argument ranges those calls see.
* ``bh_gauss_legendre`` against numpy.polynomial.legendre.leggauss (synchrotron.py:219).
* The analytic torus of ``bhgrrt.grmhd`` against the reference's own fields (golden fixture).
* Argument errors of ``bh_precompute_coefficient_bricks`` (checked before any device work).
"""

import ctypes as C
from pathlib import Path

import numpy as np
import pytest
from scipy import special

from bhgrrt import _lib
from bhgrrt import coefficient_bricks as cb
from bhgrrt.grmhd import GRMHDSnapshot, SnapshotSchemaError, generate_analytic_grmhd_torus, isco_radius

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "synchrotron_vectors.npz")


def _kv(nu, x):
    out = C.c_double()
    _lib.check(_lib.lib().bh_bessel_k(nu, x, C.byref(out)))
    return out.value


@pytest.mark.parametrize("nu", [2.0 / 3.0, 2.0, 0.0, 1.0, 5.0 / 3.0, 0.25, 2.0005, 3.5])
def test_bessel_k_matches_mpmath_and_scipy(nu):
    import mpmath

    mpmath.mp.dps = 30
    xs = np.concatenate([np.geomspace(1e-8, 1e3, 64), np.geomspace(0.5, 4.0, 41), [1.999999, 2.0, 2.000001]])
    got = np.array([_kv(nu, float(x)) for x in xs])
    exact = np.array([float(mpmath.besselk(mpmath.mpf(nu), mpmath.mpf(float(x)))) for x in xs])
    ok = exact > 1e-290   # below, the value is subnormal
    rel = np.abs(got[ok] - exact[ok]) / exact[ok]
    assert rel.max() < 1e-14, (nu, float(xs[ok][np.argmax(rel)]), rel.max())
    assert np.all(got[~ok] < 1e-280)
    # scipy (what the reference calls) is itself within ~5e-14 of the exact value near x = 2
    ref = special.kv(nu, xs)
    assert np.max(np.abs(got[ok] - ref[ok]) / ref[ok]) < 1e-13


def test_bessel_k_underflows_to_zero_like_scipy():
    for x in (1e4, 1e6, 1e8):
        assert _kv(2.0, x) == 0.0 == special.kv(2.0, x)


def test_bessel_k_rejects_bad_arguments():
    for nu, x in ((-1.0, 1.0), (2.0, 0.0), (2.0, -1.0), (float("inf"), 1.0), (float("nan"), 1.0), (1.0, float("inf"))):
        with pytest.raises(ValueError):
            _kv(nu, x)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 17, 32, 64])
def test_gauss_legendre_matches_numpy(n):
    x, w = np.zeros(n), np.zeros(n)
    _lib.check(_lib.lib().bh_gauss_legendre(n, _lib.dptr(x), _lib.dptr(w)))
    import mpmath

    mpmath.mp.dps = 40
    # exact nodes and weights: Newton on P_n in 40 digits from our nodes
    ex = [mpmath.findroot(lambda z: mpmath.legendre(n, z), mpmath.mpf(float(v))) if n > 1 or v != 0 else mpmath.mpf(0)
          for v in x]
    ew = [2 / ((1 - z * z) * mpmath.diff(lambda s: mpmath.legendre(n, s), z) ** 2) for z in ex]
    np.testing.assert_allclose(x, [float(v) for v in ex], rtol=0, atol=2e-16)
    np.testing.assert_allclose(w, [float(v) for v in ew], rtol=1e-14 if n <= 32 else 2e-13)
    # numpy's leggauss (what the reference calls) agrees to its own rounding
    rx, rw = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, rx, rtol=0, atol=4e-16)
    np.testing.assert_allclose(w, rw, rtol=5e-13 if n <= 32 else 2e-12)
    np.testing.assert_array_equal(x, -x[::-1])
    np.testing.assert_array_equal(w, w[::-1])
    # exact for polynomials of degree 2n - 1
    for k in range(0, 2 * n, 3):
        assert abs(np.sum(w * x ** k) - (2.0 / (k + 1) if k % 2 == 0 else 0.0)) < 1e-13


@pytest.mark.parametrize("tag,kw", [("a", dict(spin_a=0.5, nr=5, ntheta=4, nphi=3)),
                                    ("b", dict(spin_a=0.3, nr=6, ntheta=5, nphi=4)),
                                    ("c", dict(nr=8, ntheta=6, nphi=8))])
def test_torus_matches_reference_fields(tag, kw):
    snap = generate_analytic_grmhd_torus(**kw)
    for name in ("r", "theta", "phi", "rho", "theta_e", "pressure", "b_con"):
        np.testing.assert_allclose(getattr(snap, name), GOLD[f"{tag}_{name}"], rtol=1e-14, atol=0, err_msg=name)


def test_isco_closed_forms():
    assert isco_radius(0.0) == pytest.approx(6.0, abs=1e-12)
    assert isco_radius(0.0, prograde=False) == pytest.approx(6.0, abs=1e-12)
    assert isco_radius(0.999999, prograde=True) < 1.1
    assert isco_radius(0.999999, prograde=False) == pytest.approx(9.0, abs=1e-3)


def test_snapshot_schema_errors():
    snap = generate_analytic_grmhd_torus(nr=4, ntheta=3, nphi=3)
    with pytest.raises(SnapshotSchemaError):
        GRMHDSnapshot(snap.r[::-1], snap.theta, snap.phi, snap.rho, snap.theta_e, snap.pressure, snap.b_con, 0.5)
    with pytest.raises(SnapshotSchemaError):
        GRMHDSnapshot(snap.r, snap.theta, snap.phi, snap.rho[:, :2], snap.theta_e, snap.pressure, snap.b_con, 0.5)
    with pytest.raises(SnapshotSchemaError):
        GRMHDSnapshot(snap.r, snap.theta, snap.phi, snap.rho, snap.theta_e, snap.pressure, snap.b_con[..., :3], 0.5)
    with pytest.raises(SnapshotSchemaError):
        GRMHDSnapshot(snap.r, snap.theta, snap.phi, snap.rho, snap.theta_e, snap.pressure, snap.b_con, 1.0)


def test_from_mdot_matches_reference_scaling():
    sa = cb.PhysicalScaling.from_mdot(cb.BlackHoleSystem.m87_star(0.5), 1e-4)
    np.testing.assert_array_equal([sa.rho_cgs_per_code, sa.b_gauss_per_code, sa.mean_molecular_weight_per_electron],
                                  GOLD["a_scaling"])
    sb = cb.PhysicalScaling.from_mdot(cb.BlackHoleSystem.sgr_a_star(0.3), 1e-8)
    np.testing.assert_array_equal([sb.rho_cgs_per_code, sb.b_gauss_per_code, sb.mean_molecular_weight_per_electron],
                                  GOLD["b_scaling"])


def test_precompute_rejects_bad_arguments():
    snap = generate_analytic_grmhd_torus(nr=4, ntheta=3, nphi=3)
    sc = cb.PhysicalScaling(1.0, 1.0, 1e-18, 30.0)
    with pytest.raises(ValueError, match="stride"):
        cb.precompute_coefficient_bricks(snap, sc, stride=0)
    with pytest.raises(ValueError, match="dtype"):
        cb.precompute_coefficient_bricks(snap, sc, dtype="int8")
    lib = _lib.lib()
    one = np.zeros(4)
    out = np.zeros(11)
    bad = [cb.HybridSynchrotronCoefficients(thermal=cb.ThermalSynchrotronCoefficients(gamma_grid_size=1)),
           cb.HybridSynchrotronCoefficients(thermal=cb.ThermalSynchrotronCoefficients(pitch_nodes=17)),
           cb.HybridSynchrotronCoefficients(nonthermal=cb.NonthermalPowerLawSynchrotronCoefficients(gamma_grid_size=2000)),
           cb.HybridSynchrotronCoefficients(nonthermal=cb.NonthermalPowerLawSynchrotronCoefficients(gamma_min=0.0))]
    for model in bad:
        desc = cb.synchrotron_desc(sc, 230e9, model)
        assert lib.bh_precompute_coefficient_bricks(_lib.dptr(one), _lib.dptr(one), _lib.dptr(one), 1, C.byref(desc), 0,
                                                    out.ctypes.data_as(C.c_void_p)) == _lib.BH_ERR_VALUE
    desc = cb.synchrotron_desc(sc, float("nan"))
    assert lib.bh_precompute_coefficient_bricks(_lib.dptr(one), _lib.dptr(one), _lib.dptr(one), 1, C.byref(desc), 0,
                                                out.ctypes.data_as(C.c_void_p)) == _lib.BH_ERR_VALUE
    desc = cb.synchrotron_desc(sc, 230e9)
    assert lib.bh_precompute_coefficient_bricks(_lib.dptr(one), _lib.dptr(one), _lib.dptr(one), 1, C.byref(desc), 1,
                                                out.ctypes.data_as(C.c_void_p)) == _lib.BH_ERR_VALUE
    assert lib.bh_precompute_coefficient_bricks(_lib.dptr(one), _lib.dptr(one), _lib.dptr(one), -1, C.byref(desc), 0,
                                                out.ctypes.data_as(C.c_void_p)) == _lib.BH_ERR_VALUE
    # nothing to do: no device work, no error
    assert lib.bh_precompute_coefficient_bricks(_lib.dptr(one), _lib.dptr(one), _lib.dptr(one), 0, C.byref(desc), 0,
                                                out.ctypes.data_as(C.c_void_p)) == _lib.BH_OK


def test_estimate_brick_memory():
    est = cb.estimate_brick_memory(8, 6, 4, "float32")
    assert est["coeff_bytes"] == 8 * 6 * 4 * 11 * 4 and est["total_mib"] > 0
