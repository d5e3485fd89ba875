"""Pins the emission oracle (oracle/kgeo_oracle_emission.py) with checks
that do not re-evaluate the formula under test:

* the reference's own known-answer vectors: emissivities
  (tests/test_emissivities.py of the reference, 19 cases), bz_guess and
  bz_monopole fields (tests/test_bfields.py, agreed with ipole), the 'general'
  velocity (tests/test_velocities.py, agreed with ipole) and the 'fromfile'
  velocity and field on the reference's table a0p6_inflow.csv;
* u.u = -1 for every velocity model over random spins and radii, with the
  Kerr metric written out here (the reference's test_velocity_usq, extended
  to simfit and the drift frame);
* the tetrad is orthonormal, e_(a).e_(b) = diag(-1, 1, 1, 1);
* BZ monopole f(r): f(2) = (6 pi^2 - 49)/72 (the constant the reference's
  own field formulas use as f2) and r f(r) -> 1/4; f'(r) against a
  finite difference of f;
* the drift frame: the electric field vanishes in the fluid frame,
  u_nu F^{mu nu} = 0, with F built here from (B^i, Omega_F);
* Lambert W: w e^w = x;
* thermal synchrotron: j_nu against the Maxwell-Juettner average of the
  single-electron spectrum, sqrt(3) e^3 B sin(th) F(nu/nu_c) / (m c^2) with
  F(x) = x int_x^inf K_5/3 (Rybicki & Lightman 6.18), integrated by
  quadrature (the fit is good to 0.5 %);
* polarization: f.k = f.u = 0 and |k| in the fluid frame = 1/g (the
  redshift of calc_redshift), cos^2 2chi + sin^2 2chi = 1;
* the polarized image: the EVPA of subrings 0 and 1 against ipole's cuts in
  the reference's tests/test_images.py (three spin / inclination / velocity
  configurations, tests/golden/ipole_evpa/, the reference's tolerances).
"""

import os

import mpmath as mp
import numpy as np
import pytest
from scipy import integrate, special

from oracle import kgeo_oracle as O
from oracle import kgeo_oracle_emission as E

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TABLE = os.path.join(GOLDEN, "a0p6_inflow.csv")
MID = np.pi / 2.

# reference tests/test_emissivities.py:17-37: (model, spin, r) -> (kwargs, j)
EMIS_CASES = [
    ("constant", 0, 1, {}, 1), ("bpl", 0.3, 8, {}, 0.022089), ("bpl", 0.3, 25, {}, 0.00023715),
    ("bpl", 0.4, 3, {}, 0.36913), ("bpl", 0.5, 10, {}, 0.0085085), ("bpl", 0.9, 3.141593, {}, 0.15376),
    ("bpl", 0.3679, 3.141593, {}, 0.33512), ("glm", 0.3679, 3.141593, {}, 0.19198), ("glm", 0.5, 5, {}, 0.069270),
    ("glm", 0.3, 5, {}, 0.066282), ("glm", 0.83, 9, {}, 0.014741),
    ("glm", 0.3679, 3.1415927, dict(gamma_off=-2, sigma=1.), 0.30569),
    ("glm", 0.4, 5, dict(gamma_off=-2, sigma=1.), 0.19078), ("ring", 0.3679, 3.141593, {}, 0.061782),
    ("ring", 0.4, 5, {}, 0.75227), ("ring", 0.83, 9, {}, 0.00067951),
    ("ring", 0.5, 2.4, dict(r_ring=4, gamma_off=0.1, sigma=0.6), 0.16079),
    ("ring", 0.318, 7, dict(r_ring=4, gamma_off=0.1, sigma=0.6), 0.017807),
    ("ring", 0.1, 5., dict(r_ring=4, gamma_off=0.1, sigma=0.6), 0.32917),
]

# reference tests/test_bfields.py:15-23: (C, spin, r) -> Bphi / Br of bz_guess
BZ_GUESS_CASES = [(1, 0.5, 2.9, -0.201073), (1, 0.5, 3.7, -0.141622), (1, 0.5, 10., -0.0832697),
                  (1, 0.9375, 1.35, -382.704), (1, 0.9375, 2.8, -0.546836), (-1, 0.9375, 2.8, -0.546836)]
# reference tests/test_bfields.py:53-59: (C, spin, r) -> Br, Bth, Bph of bz_monopole
BZ_MONO_CASES = [(1, 0.1, 10., [0.009993184914347399, 0.0, -0.0001567054809727936]),
                 (1, 0.1, 2.8, [0.12737094452501346, 0.0, -0.005572446903392542]),
                 (1, 0.75, 2.35, [0.1688274613110802, 0.0, -0.07926612568926913]),
                 (-1, 0.75, 2.35, [-0.1688274613110802, 0.0, 0.07926612568926913]),
                 (1, 0.9, 1.45, [0.3990916726028977, 0.0, -11.216482663168948])]
# reference tests/test_bfields.py:33-39 and tests/test_velocities.py:64-70 (a = 0.6 table)
FROMFILE_B = [(2.15708, [0.84877, 0.0, -0.53419]), (2.5, [0.63206, 0.0, -0.23666]),
              (3.58635, [0.30706, 0.0, -0.066873]), (4.90566, [0.16411, 0.0, -0.029673])]
FROMFILE_U = [(2.15708, [5.0056, -0.54416, 0.0, 0.75899]), (2.5, [3.0016, -0.45323831, 0.0, 0.41926807]),
              (3.58635, [1.7755, -0.22549, 0.0, 0.19684]), (4.90566, [1.4859, -0.021582, 0.0, 0.12754])]
# reference tests/test_velocities.py:84-90: (retro, fac_subkep, beta_phi, beta_r, spin, r) -> u^mu
GENERAL_CASES = [((0, 1, 0.7, 0.7, 0.3, 4.0), [1.65511, -0.243725, 0, 0.141355]),
                 ((0, 0.3, 0.1, 0.4, 0.9375, 1.6), [14.0203, -0.934042, 0, 3.63511]),
                 ((0, 0.1, 0.36, 0.72, 0.1707, 3.4), [2.04774, -0.549666, 0, 0.0269644]),
                 ((1, 0.24, 0.42, 0.24, 0.42, 10.), [1.19678, -0.340182, 0, -0.00297456]),
                 ((1, 0.24, 0.42, 0.24, 0.42, 3.5), [2.23166, -0.71236, 0, 0.0121293])]

VELOCITY_MODELS = ["zamo", "infall", "kep", "cunningham", "subkep", "cunningham_subkep", "general", "gelles",
                   "simfit", "driftframe"]


def gcov(a, r, th=MID):
    """The Kerr metric in Boyer-Lindquist coordinates, written out."""
    S = r * r + a * a * np.cos(th) ** 2
    D = r * r - 2 * r + a * a
    g = np.zeros((4, 4))
    g[0, 0] = -(1 - 2 * r / S)
    g[0, 3] = g[3, 0] = -2 * a * r * np.sin(th) ** 2 / S
    g[1, 1] = S / D
    g[2, 2] = S
    g[3, 3] = (r * r + a * a + 2 * a * a * r * np.sin(th) ** 2 / S) * np.sin(th) ** 2
    return g


@pytest.mark.parametrize("case", EMIS_CASES, ids=lambda c: f"{c[0]}-{c[1]}-{c[2]}")
def test_emissivity_known_answers(case):
    model, spin, r, kw, answer = case
    j = E.Emissivity(model, **kw).jrest(spin, np.array(r))
    assert np.isclose(j, answer, rtol=1e-4)


def test_bz_guess_known_answers():
    for C, spin, r, ratio in BZ_GUESS_CASES:
        B = np.squeeze(E.Bfield("bz_guess", C=C).bfield_lab(spin, np.array([r]), th=MID))
        assert np.isclose(B[2] / B[0], ratio, rtol=1e-5)


def test_bz_monopole_known_answers_and_f():
    for C, spin, r, ref in BZ_MONO_CASES:
        B = np.squeeze(E.Bfield("bz_monopole", C=C).bfield_lab(spin, np.array([r]), th=MID))
        assert np.allclose(B, ref, rtol=1e-3, atol=1e-15)
    # f(2) is the constant f2 of the field formulas; f ~ 1/(4r)
    f2 = (6 * np.pi ** 2 - 49) / 72
    assert abs(E.bz_f(np.array([2.0]))[0] - f2) < 1e-15
    assert abs(E.bz_f(np.array([2.0 + 1e-9]))[0] - f2) < 1e-9
    r = np.array([30., 60., 99.])
    rf = r * E.bz_f(r)
    assert np.all(np.abs(rf - 0.25) < 3 / r) and np.all(np.diff(np.abs(rf - 0.25)) < 0)
    # f' against a central difference
    for x in (1.3, 2.0, 2.7, 5.0, 40.0):
        h = 1e-5 * x
        fd = (E.bz_f(np.array([x + h]))[0] - E.bz_f(np.array([x - h]))[0]) / (2 * h)
        assert abs(E.bz_fprime(np.array([x]))[0] - fd) < 1e-8 * max(1, abs(fd)) + 1e-12


def test_fromfile_models_known_answers():
    b = E.Bfield("fromfile", file=TABLE)
    for r, ref in FROMFILE_B:
        assert np.allclose(np.squeeze(b.bfield_lab(0.6, np.array([r]))), ref, rtol=1e-4)
    v = E.Velocity("fromfile", file=TABLE)
    for r, ref in FROMFILE_U:
        u = [np.squeeze(x) for x in v.u_lab(0.6, np.array([r]))]
        assert np.allclose(u, ref, rtol=1e-4)


def test_general_velocity_known_answers():
    for (retro, fk, bphi, br, spin, r), ref in GENERAL_CASES:
        v = E.Velocity("general", retrograde=bool(retro), fac_subkep=fk, beta_phi=bphi, beta_r=br)
        u = [np.squeeze(x) for x in v.u_lab(spin, np.array([r]))]
        assert np.allclose(u, ref, rtol=1e-5)


def _vel(name):
    if name == "driftframe":
        return E.Velocity("driftframe", bfield=E.Bfield("bz_monopole"), nu_parallel=0.3)
    return E.Velocity(name)


@pytest.mark.parametrize("veltype", VELOCITY_MODELS)
def test_velocity_normalisation_and_tetrad(veltype):
    rng = np.random.default_rng(42)
    v = _vel(veltype)
    eta = np.diag([-1., 1., 1., 1.])
    for spin in rng.random(3):
        reh = 1 + np.sqrt(1 - spin ** 2)
        for r in np.linspace(1.1, 15., 25):
            if r <= reh + 0.05:
                continue
            u = np.array([np.squeeze(x) for x in v.u_lab(float(spin), np.array([r]))], dtype=float)
            if not np.all(np.isfinite(u)):
                assert veltype in ("simfit", "driftframe"), (veltype, spin, r)   # no timelike solution there
                continue
            g = gcov(spin, r)
            assert np.isclose(u @ g @ u, -1., rtol=1e-9), (veltype, spin, r)
            et, ex, ey, ez = v.tetrades(float(spin), np.array([r]))
            T = np.array([[np.squeeze(c) for c in e] for e in (et, ex, ey, ez)], dtype=float)
            assert np.allclose(T @ g @ T.T, eta, atol=1e-9), (veltype, spin, r)


def test_drift_frame_has_no_electric_field():
    a = 0.8
    for bf in (E.Bfield("bz_monopole"), E.Bfield("bz_para"), E.Bfield("bz_guess"), E.Bfield("power", p=0.7)):
        v = E.Velocity("driftframe", bfield=bf, nu_parallel=-0.4)
        for r in (2.2, 3.0, 6.0, 12.0):
            for th in (MID, 1.2):
                B1, B2, B3 = [np.squeeze(x) for x in bf.bfield_lab(a, np.array([r]), th=th)]
                Om = float(np.squeeze(bf.omega_field(a, np.array([r]), th=th)))
                u = np.array([np.squeeze(x) for x in v.u_lab(a, np.array([r]), th=th)], dtype=float)
                g = gcov(a, r, th)
                # F_{mu nu} of a stationary axisymmetric field with B^i = *F^{i0}
                S = r * r + a * a * np.cos(th) ** 2
                gdet = S * np.sin(th)
                F = np.zeros((4, 4))
                F[0, 1], F[0, 2] = -gdet * Om * B2, gdet * Om * B1
                F[1, 2], F[1, 3], F[2, 3] = gdet * B3, -gdet * B2, gdet * B1
                F = F - F.T
                e = F @ u                       # e_mu = F_{mu nu} u^nu
                scale = np.abs(F).max() * np.abs(u).max()
                assert np.abs(e).max() < 1e-10 * scale, (bf.fieldtype, r, th)
                assert np.isclose(u @ g @ u, -1., rtol=1e-9)


def test_lambert_w():
    x = np.concatenate([-1 / np.e + np.logspace(-12, -1, 12), np.linspace(-0.3, 5, 40), np.logspace(1, 8, 8)])
    w = np.real(special.lambertw(x, k=0))
    assert np.allclose(w * np.exp(w), x, rtol=1e-12, atol=1e-15)


def _synchrotron_F(x):
    return x * integrate.quad(lambda y: special.kv(5. / 3., y), x, np.inf, limit=200)[0]


def test_thermal_emissivity_against_maxwell_juettner_average():
    ne, B, Te, sin_th = 1e5, 30.0, 6e10, 0.8
    Theta = E.K_B * Te / (E.M_E * E.C_LIGHT ** 2)
    nu_c0 = 1.5 * (E.E_CHARGE * B / (2 * np.pi * E.M_E * E.C_LIGHT)) * Theta ** 2 * sin_th
    P0 = np.sqrt(3) * E.E_CHARGE ** 3 * B * sin_th / (E.M_E * E.C_LIGHT ** 2)   # single electron, per Hz
    for x in (0.3, 1.0, 10.0, 100.0):
        nu = x * nu_c0
        # ultra-relativistic Maxwell-Juettner: N(gamma) = gamma^2 e^{-gamma/Theta} / (2 Theta^3); z = gamma/Theta
        avg = integrate.quad(lambda z: z * z * np.exp(-z) / 2 * _synchrotron_F(x / (z * z)), 0, 60, limit=200)[0]
        j_exact = ne * P0 * avg / (4 * np.pi)
        assert abs(E.j_nu_thermal(ne, B, Te, nu, sin_th) / j_exact - 1) < 5e-3, x


def test_polarization_vector_and_comoving_momentum():
    a, th_o = 0.6, 50 * np.pi / 180
    rng = np.random.default_rng(3)
    alpha, beta = rng.uniform(-7, 7, 400), rng.uniform(-7, 7, 400)
    lam, eta = O.conserved(a, th_o, alpha, beta)
    r_s, Ir, Imax, Nmax = O.r_equatorial(a, np.inf, th_o, 0, alpha, beta)
    ok = (Nmax >= 0) & (r_s > 1 + np.sqrt(1 - a * a) + 0.05)
    for vel, bf in ((E.Velocity("general", fac_subkep=0.7, beta_r=0.5), E.Bfield("bz_guess")),
                    (E.Velocity("kep"), E.Bfield("vert")), (E.Velocity("zamo"), E.Bfield("bz_monopole"))):
        krs = O.radial_momentum_sign(a, th_o, alpha[ok], beta[ok], Ir[ok], Imax[ok])
        kths = O.theta_momentum_sign(th_o, 0)
        r, l, e = r_s[ok], lam[ok], eta[ok]
        u = vel.u_lab(a, r)
        g = E.calc_redshift(a, r, l, e, krs, kths, *u)
        m = E.metric(a, r)
        # k^mu from the first integrals, raised with the written-out metric
        R = (r * r + a * a - a * l) ** 2 - m["Delta"] * (e + (l - a) ** 2)
        kl = np.array([-np.ones_like(r), krs * np.sqrt(R) / m["Delta"], kths * np.sqrt(e + 0 * r), l])
        et, ex, ey, ez = vel.tetrades(a, r)
        kp = np.array([sum(ei[k] * kl[k] for k in range(4)) for ei in (ex, ey, ez)])
        assert np.allclose(np.sqrt((kp ** 2).sum(0)), 1 / g, rtol=1e-9)
        sinthb, kappa, _, bsq = E.calc_polquantities(a, r, l, e, krs, kths, vel, bf)
        assert np.all((sinthb >= 0) & (sinthb <= 1 + 1e-12))
        c2, s2 = E.calc_evpa(a, th_o, alpha[ok], beta[ok], kappa)
        assert np.allclose(c2 ** 2 + s2 ** 2, 1, rtol=1e-10)
        # f = k x B / |k| in the fluid frame is orthogonal to k and to u (4D)
        bl = np.array(E.lower(m, np.array(bf.bfield_fluid(a, r, vel))))
        Bp = np.array([sum(ei[k] * bl[k] for k in range(4)) for ei in (ex, ey, ez)])
        kmag = np.sqrt((kp ** 2).sum(0))
        f3 = np.cross(kp.T, Bp.T).T / kmag
        f4 = np.array([ex[k] * f3[0] + ey[k] * f3[1] + ez[k] * f3[2] for k in range(4)])
        k4 = np.array([m["gu00"] * kl[0] + m["gu03"] * kl[3], m["gu11"] * kl[1], m["gu22"] * kl[2],
                       m["gu33"] * kl[3] + m["gu03"] * kl[0]])
        fl = np.array(E.lower(m, f4))
        fscale = np.sqrt((f3 ** 2).sum(0))
        assert np.all(np.abs((fl * k4).sum(0)) < 1e-9 * fscale * kmag)
        assert np.all(np.abs((fl * np.array([np.broadcast_to(x, r.shape) for x in u])).sum(0)) < 1e-9 * fscale *
                      np.abs(u[0]))
        assert np.allclose(sinthb, fscale / np.sqrt(bsq), rtol=1e-12)


def _evpa(Q, U):
    e = 180. / np.pi * 0.5 * np.arctan2(U, Q)
    e += 90.
    e[e > 90] -= 180.
    e[e > 90] -= 180.
    return e


# reference tests/test_images.py: (csv, spin, inclination, fac_subkep, beta_r, beta_phi, n0 limits, n1 limits,
# rtol of the four cuts)
IPOLE = [
    ("lowspin_keplerian", 0.01, 0.1, 1.0, 1.0, 1.0, [-10, -3, 3, 10], [-6.05, -5.25, 5.25, 6], (1e-2, 4e-2, 1e-2, 1e-2)),
    ("midspin_inclined_keplerian", 0.42, 60., 1.0, 1.0, 1.0, [-10, -4.3125, 5., 10.], [-5.3, -4.4, 5.7, 6.8],
     (1e-2, 3e-2, 1e-2, 1e-2)),
    ("midspin_inclined_subkep", 0.7, 45., 0.6, 0.7, 0.8, [-10, -2.7, 4, 10], [-5., -4, 5.8, 7.],
     (1e-2, 2e-2, 1e-2, 1e-2)),
]
NX = 800


def ipole_cuts(case, image_fn):
    """The reference's EVPA comparison (tests/test_images.py:21-68, 71-130):
    row 400 and column 400 of the 800 x 800, 20 M image of subrings 0 and 1
    (ring emissivity, bz_guess field, 'general' velocity, spectral index 1)
    against ipole.  image_fn(a, th_o, velocity kwargs, alpha, beta) returns
    (Q, U) of shape (npix, 2) for the given pixels."""
    name, a, inc, fk, br, bp, l0, l1, rt = case
    A, B = E.pixel_grid(-10., 10., -10., 10., 20. / NX)
    A, B = A.reshape(NX, NX), B.reshape(NX, NX)
    al = np.concatenate([A[400, :], A[:, 400]])
    be = np.concatenate([B[400, :], B[:, 400]])
    Q, U = image_fn(a, inc * np.pi / 180., dict(fac_subkep=fk, beta_phi=bp, beta_r=br), al, be)
    ev0, ev1 = _evpa(Q[:, 0], U[:, 0]), _evpa(Q[:, 1], U[:, 1])
    ploc, n0a, n0b, n1a, n1b = np.loadtxt(os.path.join(GOLDEN, "ipole_evpa", name + ".csv"), delimiter=",",
                                          unpack=True)

    def m(lim, k):
        return (ploc > lim[k]) & (ploc < lim[k + 1])

    assert np.allclose(n0a[m(l0, 0)], ev0[:NX][m(l0, 0)], rtol=rt[0])
    assert np.allclose(n0b[m(l0, 2)], ev0[NX:][m(l0, 2)], rtol=rt[1])
    assert np.allclose(n1a[m(l1, 0)], ev1[:NX][m(l1, 0)], rtol=rt[2])
    assert np.allclose(n1b[m(l1, 2)], ev1[NX:][m(l1, 2)], rtol=rt[3])


@pytest.mark.parametrize("case", IPOLE, ids=lambda c: c[0])
def test_oracle_evpa_against_ipole(case):
    def image(a, th_o, vkw, al, be):
        out = E.make_image_pixels(a, np.inf, th_o, 1, al, be, emissivity=E.Emissivity("ring"),
                                  velocity=E.Velocity("general", **vkw), bfield=E.Bfield("bz_guess", C=1.),
                                  polarization=True, specind=1.)
        return out[1], out[2]
    ipole_cuts(case, image)


def test_oracle_image_reduces_to_the_keplerian_subring():
    """With the Keplerian disk, no polarization and g^2 j, Iobs's g, r_s and
    I_r are the main oracle's equatorial_subring values (pinned there)."""
    a, th_o = 0.94, 17 * np.pi / 180
    al, be = np.meshgrid(np.linspace(-8, 8, 24), np.linspace(-8, 8, 24))
    al, be = al.ravel(), be.ravel()
    for mbar in (0, 1):
        d = E.Iobs(a, 1e4, th_o, mbar, al, be, E.Emissivity("constant"), E.Velocity("kep"), E.Bfield("rad"),
                   specind=0.)
        I, g, r_s, Ir, Imax, Nmax, zm = O.equatorial_subring(a, 1e4, th_o, mbar, al, be, p=0.0)
        assert np.array_equal(d[3], g) and np.array_equal(d[4], r_s) and np.array_equal(d[6], Ir)
        assert np.allclose(d[0], g ** 2 * (np.pi / 4), rtol=1e-15)   # j = 1 * sinthb^(1+0), sinthb = pi/4
