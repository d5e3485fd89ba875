"""Pins the oracle's geodesic samples (restatement of
kerr_raytracing_ana.py:46-683) with properties the mathematics fixes:
initial conditions, the Mino-time geodesic equations (the right-hand side of
the reference's numerical integrator, kerr_raytracing_num.py:131-153, with
tau running backwards from the screen), the equatorial consistency of
r_equatorial vs coords_at_tau (scripts/calc_t_equatorial.py:64-71), the
equatorial-crossing count against sign changes of cos(theta), and the 15
ipole known answers of tests/test_geodesics_kerr_ana.py:54-70."""

import json
import os

import numpy as np
import pytest

from oracle import kgeo_oracle as O

from .kgeo_testlib import C1, grid

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _sample_pixels(a, th, n_per=3, seed=0):
    A, B = grid(C1["window"], C1["n"])
    lam, eta = O.conserved_fixed(a, th, A, B)
    _, _, _, _, rclass = O.radial_roots(a, lam, eta)
    rng = np.random.default_rng(seed)
    idx = []
    for c in (1, 3, 4):
        w = np.where(rclass == c)[0]
        idx += list(rng.choice(w, min(n_per, len(w)), replace=False))
    w = np.where(eta < 0)[0]
    idx += list(rng.choice(w, min(n_per, len(w)), replace=False))
    return A[idx], B[idx]


def test_initial_conditions():
    a, th = C1["a"], np.deg2rad(C1["th_deg"])
    A, B = grid(C1["window"], C1["n"])
    out = O.raytrace_ana(a, C1["r_o"], th, A, B, 5)
    assert np.all(out["tau"][0] == 0)
    assert np.max(np.abs(out["r"][0] / C1["r_o"] - 1)) < 1e-10
    assert np.max(np.abs(out["th"][0] - th)) < 1e-10
    for k in ("ph", "t", "sig"):
        scale = np.maximum(np.abs(out[k]).max(axis=0), 1)
        assert np.max(np.abs(out[k][0]) / scale) < 1e-10, k


@pytest.mark.parametrize("a,th_deg", [(0.94, 17.0), (0.5, 60.0), (-0.7, 80.0), (0.9, 140.0)])
def test_mino_time_geodesic_equations(a, th_deg):
    th = np.deg2rad(th_deg)
    A, B = _sample_pixels(a, th)
    lam, eta = O.conserved_fixed(a, th, A, B)
    r1, r2, r3, r4, _ = O.radial_roots(a, lam, eta)
    tt = O.mino_total(a, C1["r_o"], eta, r1, r2, r3, r4)
    G = 4001
    tau = np.linspace(0, 0.97 * tt, G)
    h = tau[1] - tau[0]
    sig, (t, r, thv, ph) = O.coords_at_tau(a, C1["r_o"], th, A, B, tau)
    d = lambda f: (f[:-4] - 8 * f[1:-3] + 8 * f[3:-1] - f[4:]) / (12 * h)   # 4th-order centred
    rr, th2 = r[2:-2], thv[2:-2]
    Delta = rr ** 2 - 2 * rr + a ** 2
    Rp = (rr ** 2 + a ** 2 - a * lam) ** 2 - Delta * (eta + (lam - a) ** 2)
    TH = eta + (a * np.cos(th2)) ** 2 - (lam / np.tan(th2)) ** 2
    dt = (rr ** 2 + a ** 2) * (rr ** 2 + a ** 2 - a * lam) / Delta + a * (lam - a * np.sin(th2) ** 2)
    dph = a * (rr ** 2 + a ** 2 - a * lam) / Delta + lam / np.sin(th2) ** 2 - a
    dsig = rr ** 2 + (a * np.cos(th2)) ** 2
    inner = rr < 50.0   # well away from r_o, where finite differences lose digits
    for name, lhs, rhs in (("r", d(r) ** 2, Rp), ("theta", d(thv) ** 2, TH),
                           ("t", -d(t), dt), ("phi", -d(ph), dph), ("sigma", -d(sig), dsig)):
        scale = np.nanmax(np.abs(np.where(inner, rhs, 0)), axis=0, keepdims=True) + 1e-300
        err = np.where(inner, np.abs(lhs - rhs) / scale, 0)
        assert np.nanmax(err) < 1e-5, (name, np.nanmax(err))


def test_equatorial_consistency_and_crossing_count():
    a, th, r_o = 0.27118414089647946, np.deg2rad(60.0), 1.0e4
    A, B = grid(12.0, 48)
    for mbar in (0, 1, 2):
        r_s, Ir, Imax, Nmax = O.r_equatorial(a, r_o, th, mbar, A, B)
        ring = Nmax >= mbar
        assert ring.sum() > (50 if mbar == 0 else 0)
        sig, (t, r, thv, ph) = O.coords_at_tau(a, r_o, th, A[ring], B[ring], Ir[ring])
        assert np.max(np.abs(1 - thv[0] / (0.5 * np.pi))) < 1e-12
        assert np.max(np.abs(1 - r[0] / r_s[ring])) < 1e-10
    # N_max + 1 equatorial crossings along the whole ray: count sign changes of cos(theta)
    r_s, Ir, Imax, Nmax = O.r_equatorial(a, r_o, th, 0, A, B)
    rng = np.random.default_rng(2)
    idx = rng.choice(np.where(Nmax >= 0)[0], 60, replace=False)
    tau = np.linspace(0, 1.0, 20001)[:, None] * Imax[idx][None, :]
    _, (_, _, thv, _) = O.coords_at_tau(a, r_o, th, A[idx], B[idx], tau)
    cz = np.cos(thv)
    changes = np.sum(np.sign(cz[1:]) != np.sign(cz[:-1]), axis=0)
    assert np.all(changes == Nmax[idx] + 1), (changes, Nmax[idx])


def test_ipole_midplane_known_answers():
    g = json.load(open(os.path.join(GOLD, "ipole_midplane.json")))
    for spin, inc, al, be, x0, y0 in g["cases"]:
        out = O.raytrace_ana(spin, g["r_o"], inc / 180. * np.pi, np.array([al]), np.array([be]), g["ngeo"])
        r, h, p = out["r"][:, 0], out["th"][:, 0], out["ph"][:, 0]
        x = r * np.sin(h) * np.cos(p)
        y = r * np.sin(h) * np.sin(p)
        z = r * np.cos(h)

        def interp(x0_, xs, zs):
            if xs[0] > xs[1]:
                xs, zs = xs[::-1], zs[::-1]
            return np.interp(x0_, xs, zs)
        assert np.allclose([interp(0, z, x), interp(0, z, y)], [x0, y0], rtol=g["rtol"]), (spin, inc, al, be)


def test_oracle_is_stable_against_its_fixture():
    g = np.load(os.path.join(GOLD, "oracle_c1_strat.npz"))
    out = O.raytrace_ana(float(g["a"]), float(g["r_o"]), float(g["th_o"]), g["alpha"], g["beta"], int(g["ngeo"]))
    for k in ("rclass", "thclass"):
        assert np.array_equal(out[k], g[k])
    for k in ("tau", "t", "r", "th", "ph", "sig"):
        assert np.array_equal(np.isnan(out[k]), np.isnan(g[k]))
        assert np.nanmax(np.abs(out[k] - g[k]) / np.maximum(np.abs(g[k]), 1)) < 1e-12, k


def test_keplerian_velocity_normalisation_and_isco():
    # u.u = -1 with the Boyer-Lindquist metric at theta = pi/2 (tests/test_velocities.py:26-53)
    # a = 0 exactly is excluded: velocities.py:253,258 use np.sign(a) = 0 there, so
    # the reference's Keplerian branch has zero angular velocity (kept as is)
    for a in (1e-3, 0.3, 0.7, 0.94, -0.5):
        rh = 1 + np.sqrt(1 - a ** 2)
        r = np.linspace(rh + 1e-3, 15, 300)
        u0, u1, u2, u3 = O.u_kep(a, r)
        Delta = r * r - 2 * r + a * a
        g00 = -(1 - 2 / r)
        g03 = -2 * a / r
        g11 = r * r / Delta
        g33 = r * r + a * a + 2 * a * a / r
        usq = g00 * u0 ** 2 + 2 * g03 * u0 * u3 + g11 * u1 ** 2 + g33 * u3 ** 2
        assert np.max(np.abs(usq + 1)) < 1e-8, a   # Delta -> 0 at r_h + 1e-3 costs digits
    assert abs(O.r_isco(0.0) - 6.0) < 1e-14
    u0, u1, u2, u3 = O.u_kep(0.0, np.array([10.0]))
    assert u3[0] == 0.0   # the reference quirk above, reproduced
    assert abs(O.r_isco(0.9999999) - 1.0) < 0.02


def test_redshift_momentum_signs_follow_the_ray():
    # k^r and k^theta at emission point opposite to d/dtau of the backward-traced ray
    a, th, r_o = 0.5, np.deg2rad(60.0), 1.0e4
    A, B = grid(12.0, 40)
    for mbar in (0, 1):
        r_s, Ir, Imax, Nmax = O.r_equatorial(a, r_o, th, mbar, A, B)
        ok = (Nmax >= mbar) & (r_s > 1 + np.sqrt(1 - a ** 2))
        idx = np.where(ok)[0][::7]
        h = 1e-7 * Ir[idx]
        tau = np.stack([Ir[idx] - h, Ir[idx] + h])
        _, (_, r, thv, _) = O.coords_at_tau(a, r_o, th, A[idx], B[idx], tau)
        kr = -np.sign(r[1] - r[0])
        kth = -np.sign(thv[1] - thv[0])
        want_r = O.radial_momentum_sign(a, th, A[idx], B[idx], Ir[idx], Imax[idx])
        turning = np.abs(r[1] - r[0]) < 1e-9 * r_s[idx]
        assert np.all((kr == want_r) | turning)
        assert np.all(kth == O.theta_momentum_sign(th, mbar))
