"""Pins the oracle's elliptic layer (restatement of
scipy_ellip_binding.py:87-237) against mpmath and against the special
cases the reference defines."""

import os

import numpy as np
import scipy.special as sp

from oracle import kgeo_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ellip_pi_mpmath.npz")


def test_pi_matches_mpmath_including_principal_value():
    g = np.load(GOLD)
    P = O.ellip_pi(g["n"], g["phi"], g["m"])
    err = np.abs(P - g["Pi"]) / np.maximum(np.abs(g["Pi"]), 1e-3)
    # n > 1 beyond the pole is a principal value: mpmath's real part
    assert np.nanmax(err) < 1e-12, np.nanmax(err)
    assert not np.any(np.isnan(P))


def test_legendre_f_e_of_scipy_match_mpmath():
    g = np.load(GOLD)
    assert np.max(np.abs(sp.ellipkinc(g["phi"], g["m"]) - g["F"]) / np.abs(g["F"])) < 1e-13
    assert np.max(np.abs(sp.ellipeinc(g["phi"], g["m"]) - g["E"]) / np.abs(g["E"])) < 1e-13


def test_pi_special_cases_in_reference_order():
    # scipy_ellip_binding.py:125-162
    n = np.array([0.5, 0.5, 0.5, 0.5, 1.0, 2.0, np.inf, 1.0])
    phi = np.array([5e-11, 1.0, 1.0, 2.0, 2.0, np.arcsin(np.sqrt(0.5)), 1.0, 1.0])
    m = np.array([0.3, -2e14, 3.0, 1.0, 0.3, 0.3, 0.3, 0.3])
    P = O.ellip_pi(n, phi, m)
    assert P[0] == 5e-11                      # |phi| <= 1e-10 -> phi
    assert P[1] == 0.0                        # m < -1e14 -> 0
    assert np.isnan(P[2])                     # m sin^2 phi > 1 -> NaN
    assert np.isnan(P[3])                     # |phi| > pi/2 and m >= 1 -> NaN
    assert np.isinf(P[4])                     # |phi| > pi/2 and n == 1 -> inf
    rho = 1. - n[5] * np.sin(phi[5]) ** 2
    assert (rho == 0 and np.isinf(P[5])) or (rho != 0)
    assert P[6] == 0.0                        # n = inf -> 0
    assert np.isfinite(P[7])                  # n == 1 inside |phi| <= pi/2 is finite


def test_pi_periodicity_and_oddness():
    rng = np.random.default_rng(3)
    n = rng.uniform(-2, 0.9, 200)
    m = rng.uniform(-5, 0.9, 200)
    phi = rng.uniform(-1.5, 1.5, 200)
    k = rng.integers(-4, 5, 200).astype(float)
    Pc = O.ellip_pi(n, np.full(200, np.pi / 2), m)
    lhs = O.ellip_pi(n, phi + k * np.pi, m)
    rhs = 2 * k * Pc + O.ellip_pi(n, phi, m)
    assert np.max(np.abs(lhs - rhs) / np.maximum(np.abs(rhs), 1)) < 1e-13
    assert np.max(np.abs(O.ellip_pi(n, -phi, m) + O.ellip_pi(n, phi, m))) < 1e-14


def test_pi_reduces_to_f_for_n_zero():
    rng = np.random.default_rng(4)
    m = rng.uniform(-5, 0.95, 100)
    phi = rng.uniform(-9, 9, 100)
    assert np.max(np.abs(O.ellip_pi(np.zeros(100), phi, m) - sp.ellipkinc(phi, m))) < 1e-13
