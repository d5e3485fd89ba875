"""Pins the oracle's integer and fractional counts, its southern-observer
branches, region II of mino_total and the redshift composition with checks
the mathematics fixes (none of them re-evaluates the formula under test):

* n_angular_turnings (kerr_raytracing_utils.py:686-792): floor() equals the
  number of sign changes of d theta / d tau along densely sampled rays, for
  ordinary and vortical rays and both signs of beta (this includes the
  reference's vortical "norbit > 0.5" adjustment, :780-781);
* radial turning count (kerr_raytracing_utils.py:518-546, 794-826;
  equatorial_images.py:284): equals the number of sign changes of dr/dtau;
* n_poloidal_orbits (kerr_raytracing_utils.py:551-603): equals 1 at the
  Mino time where theta first returns to theta_o moving the same way, found
  by root-finding on the sampled ray;
* n_equatorial_crossings (kerr_raytracing_utils.py:605-684) at every
  intermediate sample, observers north and south of the equator: + 1 equals
  the sign changes of cos(theta) so far;
* r_equatorial for th_o > pi/2 (equatorial_lensing.py:170-180): theta(I_r)
  = pi/2 and r(I_r) = r_s;
* mino_total region II (kerr_raytracing_utils.py:485-515): mpmath.quad of
  dr / sqrt(R) from r_+ to r_o;
* calc_redshift (equatorial_images.py:307-334) with u_kep: 1 / (-k.u) with
  k^mu taken from the traced ray itself (finite differences of t, r, theta,
  phi in Mino time) and the Boyer-Lindquist metric written out; and the
  Schwarzschild circular-orbit closed form sqrt(1 - 3/r) / (1 - lam r^-3/2).
"""

import mpmath as mp
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import kgeo_oracle as O

from .kgeo_testlib import grid

R_O = 1.0e4


def _rays(a, th, window, n, keep=None):
    A, B = grid(window, n)
    lam, eta = O.conserved(a, th, A, B)
    if keep is not None:
        m = keep(eta, A, B)
        A, B = A[m], B[m]
        lam, eta = lam[m], eta[m]
    r1, r2, r3, r4, rclass = O.radial_roots(a, lam, eta)
    tt = O.mino_total(a, R_O, eta, r1, r2, r3, r4)
    return A, B, eta, rclass, tt


def _changes(x):
    """Cumulative count of sign changes of the first difference of x along
    axis 0 (the number of extrema passed), and a mask of samples within two
    samples of a change (where the dense sampling cannot resolve the order)."""
    s = np.sign(np.diff(x, axis=0))
    ch = np.zeros(x.shape, dtype=int)
    ch[2:] = np.cumsum(s[1:] != s[:-1], axis=0)
    jump = np.zeros(x.shape, dtype=bool)
    jump[1:] = ch[1:] != ch[:-1]
    near = jump.copy()
    for k in (1, 2):
        near[:-k] |= jump[k:]
        near[k:] |= jump[:-k]
    near[-2:] = True   # an extremum at the last samples has no later difference
    return ch, near


@pytest.mark.parametrize("a,th_deg", [(0.94, 17.0), (0.5, 60.0), (-0.7, 45.0), (0.3, 89.0)])
def test_angular_turnings_count_theta_extrema(a, th_deg):
    th = np.deg2rad(th_deg)
    A, B, eta, rclass, tt = _rays(a, th, 8.0, 24)
    tau = np.linspace(0, 1.0, 3001)[:, None] * (tt * O.MAXTAUFRAC)[None, :]
    _, (_, _, thv, _) = O.coords_at_tau(a, R_O, th, A, B, tau)
    nat = O.n_angular_turnings(a, th, A, B, tau)
    ch, near = _changes(thv)
    ordinary = eta > 0
    for m in (ordinary & (B <= 0), ordinary & (B > 0)):
        assert m.sum() > 100
        bad = (np.floor(nat[:, m]) != ch[:, m]) & ~near[:, m]
        assert not bad.any(), np.argwhere(bad)[:5]
    assert ch[:, ordinary].max() >= 3   # several turning points are exercised


@pytest.mark.parametrize("a,th_deg", [(0.94, 17.0), (0.99, 10.0), (0.9, 40.0), (0.99, 60.0)])
def test_vortical_turnings_count_theta_extrema(a, th_deg):
    """Vortical rays (eta < 0) live near alpha = 0, |beta| < sqrt(a^2 -
    alpha^2) cos th_o; sample them directly."""
    th = np.deg2rad(th_deg)
    rng = np.random.default_rng(3)
    al = rng.uniform(-abs(a), abs(a), 3000)
    be = rng.uniform(-1, 1, 3000) * np.sqrt(np.maximum(a * a - al * al, 0)) * np.cos(th)
    lam, eta = O.conserved(a, th, al, be)
    m = eta < 0
    A, B = al[m][:160], be[m][:160]
    lam, eta = O.conserved(a, th, A, B)
    r1, r2, r3, r4, _ = O.radial_roots(a, lam, eta)
    tt = O.mino_total(a, R_O, eta, r1, r2, r3, r4)
    tau = np.linspace(0, 1.0, 8001)[:, None] * (tt * O.MAXTAUFRAC)[None, :]
    _, (_, _, thv, _) = O.coords_at_tau(a, R_O, th, A, B, tau)
    nat = O.n_angular_turnings(a, th, A, B, tau)
    ch, near = _changes(thv)
    for sel in (B <= 0, B > 0):
        bad = (np.floor(nat[:, sel]) != ch[:, sel]) & ~near[:, sel]
        assert not bad.any(), np.argwhere(bad)[:5]
    assert ch.max() >= 1
    # the beta <= 0 adjustment (:780-781) is exercised: a ray whose formula
    # value passed 0.5 and was lifted by 0.5 ends with one turning point
    lifted = (B <= 0) & (nat[-1] > 1.0)
    if a >= 0.9:
        assert lifted.any()
        assert np.all(ch[-1, lifted] == 1)


def test_angular_turnings_rejects_southern_observer():
    # the reference accepts th_o in [0, pi/2] only (kerr_raytracing_utils.py:690-691)
    with pytest.raises(Exception, match=r"th_o should be a float in range \(0,pi/2\]"):
        O.n_angular_turnings(0.5, np.deg2rad(120.0), np.array([1.0]), np.array([1.0]), np.zeros(1))


@pytest.mark.parametrize("a,th_deg", [(0.94, 17.0), (0.5, 60.0), (-0.6, 100.0)])
def test_radial_turnings_count_r_minimum(a, th_deg):
    th = np.deg2rad(th_deg)
    A, B, eta, rclass, tt = _rays(a, th, 9.0, 26)
    tau = np.linspace(0, 1.0, 3001)[:, None] * (tt * O.MAXTAUFRAC)[None, :]
    _, (_, r, _, _) = O.coords_at_tau(a, R_O, th, A, B, tau)
    nr = O.radial_turnings(a, th, A, B, tau, tt)
    ch, near = _changes(r)
    bad = (nr != ch) & ~near
    assert not bad.any(), np.argwhere(bad)[:5]
    out = O.is_outside_crit(a, th, A, B).astype(bool)
    assert out.sum() > 100 and (~out).sum() > 50
    assert np.all(nr[-1, out] == 1) and np.all(nr[:, ~out] == 0)
    # outside the critical curve is radial class 1 (kerr_raytracing_utils.py:366-378)
    assert np.array_equal(out, rclass == 1)


def test_poloidal_orbit_period():
    a, th = 0.5, np.deg2rad(60.0)
    A, B, eta, rclass, tt = _rays(a, th, 8.0, 24, keep=lambda eta, A, B: eta > 0)
    rng = np.random.default_rng(5)
    checked = 0
    for i in rng.permutation(len(A)):
        tau = np.linspace(0, tt[i] * O.MAXTAUFRAC, 4001)
        _, (_, _, thv, _) = O.coords_at_tau(a, R_O, th, np.array([A[i]]), np.array([B[i]]), tau[:, None])
        d = thv[:, 0] - th
        s = np.sign(d[1:])
        idx = np.where(s[1:] != s[:-1])[0] + 1   # theta crosses theta_o between idx and idx + 1
        if len(idx) < 2:
            continue

        def f(x):
            _, (_, _, tq, _) = O.coords_at_tau(a, R_O, th, np.array([A[i]]), np.array([B[i]]), np.array([x]))
            return tq[0, 0] - th
        # the second return to theta_o is one full poloidal period
        period = brentq(f, tau[idx[1]], tau[idx[1] + 1], xtol=1e-15 * tau[-1], rtol=1e-15)
        npol = O.n_poloidal_orbits(a, th, np.array([A[i]]), np.array([B[i]]), np.array([period]))
        assert abs(npol[0] - 1.0) < 1e-9, (i, npol[0])
        # and it moves the same way as at the screen
        h = 1e-6 * period
        _, (_, _, t2, _) = O.coords_at_tau(a, R_O, th, np.array([A[i]]), np.array([B[i]]),
                                            np.array([[h], [period + h]]))
        assert np.sign(t2[0, 0] - th) == np.sign(t2[1, 0] - th)
        checked += 1
        if checked == 12:
            break
    assert checked == 12


@pytest.mark.parametrize("a,th_deg", [(0.5, 60.0), (0.5, 120.0), (0.94, 140.0), (-0.6, 100.0)])
def test_equatorial_crossings_at_every_sample(a, th_deg):
    th = np.deg2rad(th_deg)
    A, B, eta, rclass, tt = _rays(a, th, 10.0, 24, keep=lambda eta, A, B: eta > 0)
    tau = np.linspace(0, 1.0, 4001)[:, None] * (tt * O.MAXTAUFRAC)[None, :]
    _, (_, _, thv, _) = O.coords_at_tau(a, R_O, th, A, B, tau)
    neq = O.n_equatorial_crossings(a, th, A, B, tau)
    s = np.sign(np.cos(thv))
    ch = np.zeros(thv.shape, dtype=int)
    ch[1:] = np.cumsum(s[1:] != s[:-1], axis=0)
    jump = np.zeros(thv.shape, dtype=bool)
    jump[1:] = ch[1:] != ch[:-1]
    near = jump.copy()
    near[:-1] |= jump[1:]
    near[1:] |= jump[:-1]
    for m in (B <= 0, B > 0):
        bad = (neq[:, m] + 1 != ch[:, m]) & ~near[:, m]
        assert not bad.any(), np.argwhere(bad)[:5]
    assert ch.max() >= 3


@pytest.mark.parametrize("th_deg", [120.0, 140.0])
def test_r_equatorial_southern_observer(th_deg):
    a, th = 0.94, np.deg2rad(th_deg)
    A, B = grid(10.0, 40)
    for mbar in (0, 1, 2):
        r_s, Ir, Imax, Nmax = O.r_equatorial(a, R_O, th, mbar, A, B)
        ring = Nmax >= mbar
        assert ring.sum() > (20 if mbar < 2 else 5)
        _, (_, r, thv, _) = O.coords_at_tau(a, R_O, th, A[ring], B[ring], Ir[ring])
        assert np.max(np.abs(thv[0] - 0.5 * np.pi)) < 1e-13
        assert np.max(np.abs(r[0] / r_s[ring] - 1)) < 1e-10
        # I_r is the (mbar+1)-th crossing: cos(theta) changes sign mbar times before it
        idx = np.where(ring)[0][::max(1, ring.sum() // 25)]
        tau = np.linspace(0, 1.0 - 1e-9, 4001)[:, None] * Ir[idx][None, :]
        _, (_, _, tq, _) = O.coords_at_tau(a, R_O, th, A[idx], B[idx], tau)
        s = np.sign(np.cos(tq))
        assert np.all(np.sum(s[1:] != s[:-1], axis=0) == mbar)


def test_mino_total_region_two_matches_quadrature():
    """Horizon-bound rays with four real roots (rclass 2), which the image of
    a fast retrograde hole seen edge-on holds (50 of 128^2 pixels here):
    tau_tot = int_{r_+}^{r_o} dr/sqrt(R)."""
    a, th = -0.99, np.deg2rad(89.0)
    A, B = grid(6.0, 128)
    lam, eta = O.conserved_fixed(a, th, A, B)
    r1, r2, r3, r4, rclass = O.radial_roots(a, lam, eta)
    tt = O.mino_total(a, R_O, eta, r1, r2, r3, r4)
    rplus = 1 + np.sqrt(1 - a ** 2)
    two = np.where((rclass == 2) & (np.real(r3) < rplus))[0]
    assert len(two) > 20
    mp.mp.dps = 30
    for i in np.random.default_rng(7).choice(two, 6, replace=False):
        rts = [mp.mpf(float(np.real(x[i]))) for x in (r1, r2, r3, r4)]
        Rf = lambda r: mp.fprod([r - x for x in rts])
        val = mp.quad(lambda r: 1 / mp.sqrt(Rf(r)), [rplus, 3, 10, 100, R_O])
        assert abs(tt[i] - float(val)) / float(val) < 1e-12, (i, tt[i], float(val))


def _kerr_metric_eq(a, r):
    """Boyer-Lindquist metric components at theta = pi/2 (Sigma = r^2)."""
    S = r * r
    D = r * r - 2 * r + a * a
    gtt = -(1 - 2 * r / S)
    gtp = -2 * a * r / S
    gpp = r * r + a * a + 2 * a * a * r / S
    grr = S / D
    gthth = S
    return gtt, gtp, gpp, grr, gthth


def test_redshift_is_minus_inverse_k_dot_u():
    """g from calc_redshift with u_kep and the reference's momentum signs
    equals 1 / (-g_{mu nu} k^mu u^nu), k^mu = -(1/Sigma) dx^mu/dtau taken from
    the backward-traced ray by finite differences at the emission point."""
    for a, thd in ((0.5, 60.0), (-0.7, 40.0), (0.94, 17.0)):
        th = np.deg2rad(thd)
        A, B = grid(12.0, 40)
        lam, eta = O.conserved(a, th, A, B)
        rh = 1 + np.sqrt(1 - a ** 2)
        checked_plunge = 0
        for mbar in (0, 1):
            r_s, Ir, Imax, Nmax = O.r_equatorial(a, R_O, th, mbar, A, B)
            # emission within 100 M: far out, dr/dtau ~ r^2 and the fixed
            # Mino-time stencil below would not resolve the ray
            ok = (Nmax >= mbar) & (r_s > rh * 1.05) & (r_s < 100.0)
            idx = np.where(ok)[0][::5]
            krs = O.radial_momentum_sign(a, th, A[idx], B[idx], Ir[idx], Imax[idx])
            kths = O.theta_momentum_sign(th, mbar)
            u0, u1, u2, u3 = O.u_kep(a, r_s[idx])
            g = O.calc_redshift(a, r_s[idx], lam[idx], eta[idx], krs, kths, u0, u1, u2, u3)
            h = 1e-4 * np.minimum(Ir[idx], 1.0)
            steps = np.array([-2, -1, 1, 2])[:, None]
            tau = Ir[idx][None, :] + steps * h[None, :]
            _, (t, r, thv, ph) = O.coords_at_tau(a, R_O, th, A[idx], B[idx], tau)
            d = lambda f: (f[0] - 8 * f[1] + 8 * f[2] - f[3]) / (12 * h)
            S = r_s[idx] ** 2
            kt, kr, kph = -d(t) / S, -d(r) / S, -d(ph) / S
            gtt, gtp, gpp, grr, _ = _kerr_metric_eq(a, r_s[idx])
            # u^theta = 0 for the equatorial disk
            kdotu = gtt * kt * u0 + gtp * (kt * u3 + kph * u0) + gpp * kph * u3 + grr * kr * u1
            g_ind = 1.0 / (-kdotu)
            # away from the radial turning point, where dr/dtau -> 0 and the
            # finite difference of r loses its digits
            R = (r_s[idx] ** 2 + a * a - a * lam[idx]) ** 2 - (r_s[idx] ** 2 - 2 * r_s[idx] + a * a) * (
                eta[idx] + (lam[idx] - a) ** 2)
            fine = R > 1e-3 * r_s[idx] ** 4
            assert fine.sum() > 20
            err = np.abs(g_ind / g - 1)[fine]
            assert np.max(err) < 1e-7, (a, thd, mbar, np.max(err))
            checked_plunge += int(np.sum(fine & (u1 != 0)))
        assert checked_plunge > 0   # the plunging region (u^r != 0) is exercised


def test_redshift_schwarzschild_circular_orbits():
    """a -> 0 outside the ISCO: g = sqrt(1 - 3/r) / (1 - lam r^-3/2)."""
    a = 1e-9   # velocities.py:253,258 need sign(a) != 0 for the orbital frequency
    r = np.linspace(6.0, 40.0, 57)
    lam = np.linspace(-5.0, 5.0, 57)
    eta = np.full(57, 20.0)
    u0, u1, u2, u3 = O.u_kep(a, r)
    assert np.all(u1 == 0)
    g = O.calc_redshift(a, r, lam, eta, np.ones(57), np.ones(57), u0, u1, u2, u3)
    want = np.sqrt(1 - 3 / r) / (1 - lam * r ** -1.5)
    assert np.max(np.abs(g / want - 1)) < 1e-7
