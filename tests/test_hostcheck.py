"""The device headers (kg_math.hpp, kg_geodesic.hpp) compiled for the CPU
(libkgeo_hostcheck.so) against mpmath truth and the oracle.  This checks the
exact source the HIP kernels are built from, without a GPU."""

import os

import numpy as np
import pytest
import scipy.special as sp

from oracle import kgeo_oracle as O

from .kgeo_testlib import C1, Host, c2_subset, conditioning_mask, errors, grid, summarize

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ellip_pi_mpmath.npz")


@pytest.fixture(scope="module")
def host():
    return Host()


def test_carlson_against_scipy(host):
    rng = np.random.default_rng(0)
    n = 20000
    x, y = rng.uniform(0, 5, n), rng.uniform(0, 5, n)
    z, p = rng.uniform(1e-3, 5, n), rng.uniform(1e-3, 5, n)
    x[:100] = 0
    rel = lambda a, b: np.max(np.abs(a - b) / np.abs(b))
    assert rel(host.rf(x, y, z), sp.elliprf(x, y, z)) < 2e-15
    assert rel(host.rd(x, y, z), sp.elliprd(x, y, z)) < 3e-15
    assert rel(host.rj(x, y, z, p), sp.elliprj(x, y, z, p)) < 3e-15
    xc, yc = rng.uniform(0, 5, n), rng.uniform(-5, 5, n)
    assert rel(host.rc(xc, yc), sp.elliprc(xc, yc)) < 2e-15
    yc2 = xc * (1 + rng.uniform(-3e-3, 3e-3, n))
    assert rel(host.rc(xc, yc2), sp.elliprc(xc, yc2)) < 2e-15


@pytest.mark.parametrize("a,th_deg", [(0.94, 17.0), (0.0, 60.0), (5e-5, 45.0), (-0.9, 80.0), (0.7, 140.0)])
def test_plan_halves_equal_whole_plan(host, a, th_deg):
    """The plan kernels' split (kg_setup_theta, then kg_setup_radial on the
    side stream) writes every plan slot, label, tau_tot, nmax_eq and status
    bit for bit as the whole-plan kernel kg_setup does."""
    import ctypes
    A, B = grid(9.0, 40)
    A = np.concatenate([A, [0.0, 1e-9, 3.0]])
    B = np.concatenate([B, [5.0, 0.0, 0.0]])
    sc = host.scalars(a, 1.0e4, np.deg2rad(th_deg))
    fn = host.lib.kgh_plan_split_mismatches
    fn.restype = ctypes.c_int64
    bad = fn(host._p(sc), host._p(np.ascontiguousarray(A)), host._p(np.ascontiguousarray(B)),
             ctypes.c_int64(len(A)), ctypes.c_int64(500))
    assert bad == 0


def test_acos_poly_against_mpmath(host):
    """The theta outputs' arccos (kg_math.hpp acos_poly: half-angle reduction
    and a degree-11 polynomial for asin) within 1 ulp of the exact value, as
    numpy's arccos, on both sides of the |x| = 1/2 switch, next to +-1 and 0,
    and NaN outside [-1, 1]."""
    import mpmath as mp
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-1, 1, 4000), 1 - 10.0 ** rng.uniform(-16, 0, 500),
                        -1 + 10.0 ** rng.uniform(-16, 0, 500), 0.5 + rng.uniform(-1e-12, 1e-12, 200),
                        -0.5 + rng.uniform(-1e-12, 1e-12, 200), 10.0 ** rng.uniform(-300, 0, 300),
                        [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, np.nextafter(0.5, 0), np.nextafter(-0.5, 0)]])
    got, _ = host.special(3, x)
    with mp.workdps(40):
        ref = np.array([float(mp.acos(mp.mpf(v))) for v in x])
    ulp = np.spacing(np.abs(ref))
    assert np.max(np.abs(got - ref) / ulp) <= 1.0
    assert got[x == 1.0][0] == 0.0 and got[x == -1.0][0] == np.pi
    bad, _ = host.special(3, np.array([1.0 + 2e-16, -1.5, np.nan, np.inf]))
    assert np.isnan(bad).all()


def test_rj_principal_value_matches_transform(host):
    # the reference's DLMF 19.20.14 transform (scipy_ellip_binding.py:224-230) with z = 1
    rng = np.random.default_rng(1)
    n = 5000
    x, y = rng.uniform(0, 1, n), rng.uniform(0, 3, n)
    z = np.ones(n)
    q = rng.uniform(1e-3, 5, n)
    pp = (z * (x + y + q) - x * y) / (z + q)
    want = ((pp - z) * sp.elliprj(x, y, z, pp) - 3 * sp.elliprf(x, y, z)
            + 3 * np.sqrt(x * y * z / (x * y + pp * q)) * sp.elliprc(x * y + pp * q, pp * q)) / (q + z)
    got = host.rj(x, y, z, -q)
    # the transform cancels O(1) terms near the zero of R_J: absolute scale 1
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1.0)) < 1e-14


def test_jacobi_against_scipy(host):
    rng = np.random.default_rng(2)
    n = 20000
    u = rng.uniform(-30, 30, n)
    m = rng.uniform(0, 1 - 1e-9, n)   # scipy's m > 1-1e-9 branch is inexact for |u| > K
    m[:30] = 10 ** rng.uniform(-15, -3, 30)
    m[30:40] = 0
    sn, cn, dn, K = host.ellipj(u, m)
    ssn, scn, sdn, _ = sp.ellipj(u, m)
    assert np.max(np.abs(sn - ssn)) < 1e-13
    assert np.max(np.abs(cn - scn)) < 1e-13
    assert np.max(np.abs(dn - sdn) / sdn) < 1e-11
    assert np.max(np.abs(K - sp.ellipk(m)) / sp.ellipk(m)) < 1e-15


def test_legendre_against_mpmath(host):
    g = np.load(GOLD)
    F, E, P = host.legendre(g["n"], g["phi"], g["m"])
    assert np.max(np.abs(F - g["F"]) / np.abs(g["F"])) < 1e-13
    assert np.max(np.abs(E - g["E"]) / np.abs(g["E"])) < 1e-13
    assert np.max(np.abs(P - g["Pi"]) / np.maximum(np.abs(g["Pi"]), 1e-3)) < 1e-12


def test_pi_special_cases_match_oracle(host):
    n = np.array([0.5, 0.5, 0.5, 0.5, 1.0, np.inf, 1.0, 2.5, 0.3])
    phi = np.array([5e-11, 1.0, 1.0, 2.0, 2.0, 1.0, 1.0, 7.0, -40.0])
    m = np.array([0.3, -2e14, 3.0, 1.0, 0.3, 0.3, 0.3, 0.7, -3.0])
    _, _, P = host.legendre(n, phi, m)
    want = O.ellip_pi(n, phi, m)
    assert np.array_equal(np.isnan(P), np.isnan(want))
    assert np.array_equal(np.isinf(P), np.isinf(want))
    ok = np.isfinite(want)
    assert np.max(np.abs(P[ok] - want[ok]) / np.maximum(np.abs(want[ok]), 1)) < 1e-13


@pytest.mark.parametrize("case", [
    (0.94, 17.0, 8.0, 64, 100),
    (0.27118414089647946, 60.0, 12.0, 96, 60),
    (-0.9, 80.0, 12.0, 96, 60),
    (0.0, 30.0, 12.0, 48, 40),
    (0.99, 5.0, 12.0, 96, 60),
])
def test_device_source_matches_oracle(host, case):
    a, th_deg, W, n, G = case
    th = np.deg2rad(th_deg)
    A, B = grid(W, n)
    ref = O.raytrace_ana(a, C1["r_o"], th, A, B, G)
    got = host.raytrace(a, C1["r_o"], th, A, B, G)
    setup = host.setup(a, C1["r_o"], th, A, B)
    assert np.array_equal(setup["rclass"], ref["rclass"])
    assert np.array_equal(setup["thclass"], ref["thclass"])
    nm = O.n_equatorial_crossings(a, th, A, B, ref["tau"][-1])
    assert np.array_equal(got["nmax_eq"], nm)
    assert np.max(np.abs(got["tau"] - ref["tau"]) / np.maximum(ref["tau"], 1e-300)) < 1e-12
    ill = conditioning_mask(a, C1["r_o"], th, A, B, ref["tau"], ref)
    for k in ("r", "th", "ph", "t", "sig"):
        s = summarize(got[k], ref[k], k, ill)
        assert s["nan_mismatch"] == 0, (k, s)
        assert s["max_well"] < 1e-10, (k, s)
        assert s["p99"] < 1e-10, (k, s)


def test_device_source_c2_subset(host):
    A, B, _ = c2_subset()
    A, B = A[:1024], B[:1024]
    a, th = 0.94, np.deg2rad(17.0)
    ref = O.raytrace_ana(a, 1e4, th, A, B, 500)
    got = host.raytrace(a, 1e4, th, A, B, 500)
    ill = conditioning_mask(a, 1e4, th, A, B, ref["tau"], ref)
    for k in ("r", "th", "ph", "t", "sig"):
        s = summarize(got[k], ref[k], k, ill)
        assert s["max_well"] < 1e-10 and s["p99"] < 1e-10, (k, s)
