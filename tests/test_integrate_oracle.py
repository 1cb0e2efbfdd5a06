"""Pins of the FP64 integrator oracle (oracle/integrate_oracle.py) and the
library's host-side cosmology against it.  No GPU needed.

The oracle is checked against:
* the reference's own tests: EdS closed forms for advance_a and the
  drift/kick integrals (crates/gadget-ng-core/src/cosmology.rs:784-882,
  same tolerances), flat space giving dt, dt/2, dt/2 (:872-882), and the flat
  cosmological step equal to the Newtonian one
  (crates/gadget-ng-integrators/src/leapfrog.rs:103-124);
* an independent integration of the Friedmann equation (scipy, with the dark
  energy density from its continuity equation rather than the closed form);
* closed forms of the KDK map (constant force; one oscillator step) and its
  time reversibility;
* the edge cases of rem_euclid.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad, solve_ivp

import integrate_oracle as io
import gadget_ng_pm_hip as pm
from gadget_ng_pm_hip import integrate as gi

EDS = (1.0, 0.0, 1.0, -1.0, 0.0, 0.0)


def eds_a(a0, h0, t):
    return (a0 ** 1.5 + 1.5 * h0 * t) ** (2.0 / 3.0)


def test_advance_a_eds_matches_analytic():
    a = io.advance_a(EDS, 1.0, 0.01)
    assert abs(a - eds_a(1.0, 1.0, 0.01)) / eds_a(1.0, 1.0, 0.01) < 1e-8
    a = 1.0
    for _ in range(10):
        a = io.advance_a(EDS, a, 0.01)
    assert abs(a - eds_a(1.0, 1.0, 0.1)) / eds_a(1.0, 1.0, 0.1) < 1e-6


def test_drift_kick_factors_eds_match_analytic():
    dt = 0.05
    drift, kh, kh2 = io.drift_kick_factors(EDS, 1.0, dt)
    a_end = eds_a(1.0, 1.0, dt)
    a_mid = eds_a(1.0, 1.0, dt / 2)
    drift_ana = 2.0 * (1.0 - math.sqrt(1.0 / a_end))
    kh_ana = 2.0 * (math.sqrt(a_mid) - 1.0)
    kh2_ana = 2.0 * (math.sqrt(eds_a(a_mid, 1.0, dt / 2)) - math.sqrt(a_mid))
    assert abs(drift - drift_ana) / drift_ana < 1e-7
    assert abs(kh - kh_ana) / kh_ana < 1e-7
    assert abs(kh2 - kh2_ana) / kh2_ana < 1e-7


def test_flat_space_factors_are_dt():
    flat = (0.0, 0.0, 0.0, -1.0, 0.0, 0.0)
    d, k1, k2 = io.drift_kick_factors(flat, 1.0, 0.1)
    assert abs(d - 0.1) < 1e-12 and abs(k1 - 0.05) < 1e-12 and abs(k2 - 0.05) < 1e-12


@pytest.mark.parametrize("c", [(0.3, 0.7, 0.1, -1.0, 0.0, 0.0), (0.3, 0.7, 0.1, -0.8, 0.3, 0.0),
                               (0.28, 0.7, 0.2, -0.9, -0.2, 0.02)])
def test_advance_a_and_factors_against_an_independent_friedmann_integration(c):
    """a(t) from scipy's RK45 on da/dt = a h0 sqrt((Om+Onu)/a^3 + rho_de(a)),
    with rho_de from d ln rho/d ln a = -3(1 + w0 + wa(1 - a)) integrated by quad."""
    om, ol, h0, w0, wa, onu = c

    def rho_de(a):
        if a == 1.0:
            return ol
        val, _ = quad(lambda x: -3.0 * (1.0 + w0 + wa * (1.0 - x)) / x, 1.0, a, epsabs=1e-14, epsrel=1e-13)
        return ol * math.exp(val)

    def rhs(t, y):
        a = y[0]
        return [a * h0 * math.sqrt((om + onu) / a ** 3 + rho_de(a))]

    a0, dt = 0.5, 0.4
    sol = solve_ivp(rhs, (0.0, dt), [a0], rtol=1e-12, atol=1e-14, dense_output=True)
    assert abs(io.advance_a(c, a0, dt) - sol.y[0, -1]) < 1e-7
    drift, kh, kh2 = io.drift_kick_factors(c, a0, dt)
    a_of = lambda t: sol.sol(t)[0]
    d_ref = quad(lambda t: 1.0 / a_of(t) ** 2, 0.0, dt, epsabs=1e-13)[0]
    k1_ref = quad(lambda t: 1.0 / a_of(t), 0.0, dt / 2, epsabs=1e-13)[0]
    k2_ref = quad(lambda t: 1.0 / a_of(t), dt / 2, dt, epsabs=1e-13)[0]
    assert abs(drift - d_ref) / d_ref < 1e-8
    assert abs(kh - k1_ref) / k1_ref < 1e-8
    assert abs(kh2 - k2_ref) / k2_ref < 1e-8


def test_kdk_is_exact_for_a_constant_force():
    rng = np.random.default_rng(3)
    x0, v0, acc = rng.normal(size=(5, 3)), rng.normal(size=(5, 3)), rng.normal(size=(5, 3))
    dt, n = 0.125, 8
    x, v = x0, v0
    for _ in range(n):
        x, v, _ = io.kdk_step(x, v, (dt, dt / 2, dt / 2), lambda p: acc)
    t = dt * n
    np.testing.assert_allclose(x, x0 + v0 * t + 0.5 * acc * t * t, rtol=0, atol=1e-13)
    np.testing.assert_allclose(v, v0 + acc * t, rtol=0, atol=1e-13)


def test_one_oscillator_step_closed_form():
    dt = 0.1
    x, v, a = io.kdk_step(np.array([[1.0, 0, 0]]), np.zeros((1, 3)), (dt, dt / 2, dt / 2), lambda p: -p)
    assert abs(x[0, 0] - (1.0 - dt * dt / 2)) < 1e-15
    assert abs(v[0, 0] - (-dt + dt ** 3 / 4)) < 1e-15
    assert abs(a[0, 0] + x[0, 0]) == 0.0


def test_kdk_is_time_reversible():
    rng = np.random.default_rng(4)
    x0, v0 = rng.normal(size=(6, 3)), rng.normal(size=(6, 3))
    force = lambda p: -p * (1.0 + 0.1 * (p ** 2).sum(axis=1, keepdims=True))
    x, v = x0, v0
    for _ in range(20):
        x, v, _ = io.kdk_step(x, v, (0.05, 0.025, 0.025), force)
    for _ in range(20):
        x, v, _ = io.kdk_step(x, v, (-0.05, -0.025, -0.025), force)
    np.testing.assert_allclose(x, x0, atol=1e-12)
    np.testing.assert_allclose(v, v0, atol=1e-12)


def test_cosmological_step_with_flat_factors_is_the_newtonian_step():
    force = lambda p: -p
    x0, v0 = np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 0.5, 0.0]])
    dt = 0.05
    a = io.kdk_step(x0, v0, (dt, 0.5 * dt, 0.5 * dt), force)
    cf = gi.CosmoFactors.flat(dt)
    b = io.kdk_step(x0, v0, (cf.drift, cf.kick_half, cf.kick_half2), force)
    for u, w in zip(a, b):
        assert np.array_equal(u, w)


def test_rem_euclid_edge_cases():
    L = 256.0
    x = np.array([-0.5, 0.0, L, 2.5 * L, -3 * L, 17.25, -1e-20])
    r = io.rem_euclid(x, L)
    assert list(r[:6]) == [L - 0.5, 0.0, 0.0, 0.5 * L, 0.0, 17.25]
    # the reference's own rounding edge: a tiny negative lands on L itself
    assert r[6] == L


@pytest.mark.parametrize("c", [(1.0, 0.0, 1.0, -1.0, 0.0, 0.0), (0.3, 0.7, 0.1, -1.0, 0.0, 0.0),
                               (0.3, 0.7, 0.1, -0.8, 0.3, 0.0), (0.28, 0.7, 0.2, -0.9, -0.2, 0.02),
                               (0.0, 0.0, 0.0, -1.0, 0.0, 0.0)])
def test_library_cosmology_equals_the_oracle_bit_for_bit(c):
    cp = gi.CosmologyParams(*c)
    for a0, dt in ((1.0, 0.05), (0.02, 0.001), (0.5, 0.4)):
        assert cp.advance_a(a0, dt) == io.advance_a(c, a0, dt)
        assert cp.drift_kick_factors(a0, dt) == io.drift_kick_factors(c, a0, dt)


def test_library_rejects_bad_integrator_arguments():
    import ctypes
    L = pm.lib()
    h = ctypes.c_void_p()
    assert L.gng_kdk_create(None, 10, ctypes.byref(h)) == pm.GNG_ERR_INVALID
    assert L.gng_kdk_step(None, 1.0, 0.5, 0.5, 1.0, 0, None) == pm.GNG_ERR_INVALID
    assert L.gng_kdk_run(None, None, None, 1.0, 1.0, 1, 0, None) == pm.GNG_ERR_INVALID
    assert L.gng_cosmo_drift_kick_factors(None, 1.0, 0.1, None) == pm.GNG_ERR_INVALID
    assert math.isnan(L.gng_cosmo_advance_a(None, 1.0, 0.1))
