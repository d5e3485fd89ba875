"""The count kernels' source (kg_counts.hpp) on the CPU against the oracle:
equatorial crossings, angular turning points (floor of n_angular_turnings,
kerr_raytracing_utils.py:686-792) and the radial turning count bit-exact,
n_angular_turnings and n_poloidal_orbits within 1e-12, on the C1 image at
every sample and on other spins and inclinations (vortical rays, a southern
observer for the counts that allow one)."""

import numpy as np
import pytest

from oracle import kgeo_oracle as O

from .kgeo_testlib import C1, Host, grid


def _case(a, th, A, B, G, r_o=1.0e4):
    lam, eta = O.conserved_fixed(a, th, A, B)
    r1, r2, r3, r4, _ = O.radial_roots(a, lam, eta)
    tt = O.mino_total(a, r_o, eta, r1, r2, r3, r4)
    tau = O.mino_grid(tt, G)
    return tau, tt


def _close(x, y, tol):
    assert np.array_equal(np.isnan(x), np.isnan(y))
    ok = ~np.isnan(y)
    return np.max(np.abs(x[ok] - y[ok]) / np.maximum(np.abs(y[ok]), 1.0), initial=0.0) < tol


@pytest.mark.parametrize("a,th_deg,window,n,G", [
    (C1["a"], C1["th_deg"], C1["window"], C1["n"], C1["ngeo"]),
    (0.99, 10.0, 3.0, 48, 80),       # many vortical rays
    (-0.7, 60.0, 9.0, 40, 60),
    (0.0, 30.0, 8.0, 30, 40),        # |a| < MINSPIN
    (0.3, 89.0, 9.0, 30, 40),
])
def test_counts_match_oracle(a, th_deg, window, n, G):
    th = np.deg2rad(th_deg)
    A, B = grid(window, n)
    tau, tt = _case(a, th, A, B, G)
    got = Host().counts(a, 1.0e4, th, A, B, tau)
    assert np.array_equal(got["n_equatorial"], O.n_equatorial_crossings(a, th, A, B, tau))
    assert _close(got["n_poloidal"], O.n_poloidal_orbits(a, th, A, B, tau), 1e-12)
    nat = O.n_angular_turnings(a, th, A, B, tau)
    assert _close(got["n_angular"], nat, 1e-12)
    assert np.array_equal(got["turns_theta"], np.floor(nat).astype(np.int32))
    assert np.array_equal(got["turns_r"], O.radial_turnings(a, th, A, B, tau, tt))
    assert got["turns_theta"].max() >= 1
    if window >= 8.0:   # the image reaches outside the critical curve
        assert got["turns_r"].max() == 1


def test_counts_southern_observer():
    """Equatorial crossings and the radial count hold for th_o > pi/2 too
    (the reference's n_angular_turnings refuses such observers)."""
    a, th = 0.94, np.deg2rad(140.0)
    A, B = grid(10.0, 40)
    tau, tt = _case(a, th, A, B, 60)
    got = Host().counts(a, 1.0e4, th, A, B, tau)
    assert np.array_equal(got["n_equatorial"], O.n_equatorial_crossings(a, th, A, B, tau))
    assert np.array_equal(got["turns_r"], O.radial_turnings(a, th, A, B, tau, tt))
