"""Pins the oracle's lensed-curve functions (equatorial_lensing.py:198-293)
with what the mathematics fixes:

* rho_of_req returns roots: r_equatorial2 - req changes sign across each
  returned rho (the objective is the pinned r_equatorial, with the
  reference's 1e50 / 1 mapping of rays without the crossing);
* a = 0 has no frame dragging, so every curve is symmetric under
  alpha -> -alpha: rho(varphi) = rho(pi - varphi);
* critical_curve: the conserved quantities of its parametrization give R(r)
  a double root at the photon-orbit radius r (R(r) = R'(r) = 0), and its
  points separate rays inside the critical curve (is_outside_crit 0) from
  rays outside it (1).
"""

import numpy as np
import pytest

from oracle import kgeo_oracle as O


@pytest.mark.parametrize("a,th_deg,req,mbar", [(0.5, 30.0, 6.0, 0), (0.9, 60.0, None, 1), (-0.6, 75.0, 8.0, 2)])
def test_rho_of_req_returns_roots(a, th_deg, req, mbar):
    th = np.deg2rad(th_deg)
    req = (1 + np.sqrt(1 - a * a)) if req is None else req
    vp = np.linspace(-180, 179, 24) * np.pi / 180
    _, rhos, al, be = O.rho_of_req(a, th, req, mbar, vp)
    assert np.allclose(al, np.cos(vp) * rhos) and np.allclose(be, np.sin(vp) * rhos)
    for v, r in zip(vp, rhos):
        lo, hi = (O.r_equatorial2(r * (1 - s), v, a, th, mbar) - req for s in (1e-9, -1e-9))
        mid = O.r_equatorial2(r, v, a, th, mbar) - req
        assert lo * hi <= 0 or abs(mid) < 1e-9 * req, (v, r, lo, mid, hi)


def test_schwarzschild_curves_are_mirror_symmetric():
    a, th = 0.0, np.deg2rad(50.0)
    vp = np.array([10.0, 35.0, 80.0, 125.0, -40.0, -100.0]) * np.pi / 180
    for req, mbar in ((6.0, 0), (4.0, 1)):
        _, r1, _, _ = O.rho_of_req(a, th, req, mbar, vp)
        _, r2, _, _ = O.rho_of_req(a, th, req, mbar, np.pi - vp)
        assert np.max(np.abs(r1 - r2)) < 1e-9, (req, mbar, r1, r2)
    # spin breaks it
    _, r1, _, _ = O.rho_of_req(0.9, th, 6.0, 0, vp)
    _, r2, _, _ = O.rho_of_req(0.9, th, 6.0, 0, np.pi - vp)
    assert np.max(np.abs(r1 - r2)) > 1e-3


@pytest.mark.parametrize("a,th_deg", [(0.5, 30.0), (0.94, 17.0), (-0.7, 70.0)])
def test_critical_curve_is_the_photon_orbit_boundary(a, th_deg):
    th = np.deg2rad(th_deg)
    al, be = O.critical_curve(a, th, n=2000)
    half = len(al) // 2
    al, be = al[:half], be[:half]
    assert len(al) > 200 and np.all(be >= 0)
    lam, eta = O.conserved(a, th, al, be)
    rminus = 2 * (1 + np.cos(2. * np.arccos(-a) / 3.))
    rplus = 2 * (1 + np.cos(2. * np.arccos(a) / 3.))
    r = np.linspace(rminus, rplus, 2000)
    lam_p = -(r ** 3 + a ** 2 + r * a ** 2 - 3 * r ** 2) / (a * r - a)
    eta_p = r ** 2 * (3 * r ** 2 + a ** 2 - lam_p ** 2) / (r ** 2 - a ** 2)
    ok = np.isfinite(-lam_p / np.sin(th)) & np.isfinite(np.sqrt(eta_p + (a * np.cos(th)) ** 2 - (lam_p / np.tan(th)) ** 2))
    r, lam_p, eta_p = r[ok], lam_p[ok], eta_p[ok]
    assert np.allclose(lam, lam_p, rtol=1e-12, atol=1e-12)
    # R(r) = (r^2 + a^2 - a lam)^2 - Delta (eta + (lam - a)^2) has a double root at r
    R = (r ** 2 + a ** 2 - a * lam_p) ** 2 - (r ** 2 - 2 * r + a ** 2) * (eta_p + (lam_p - a) ** 2)
    dR = 4 * r * (r ** 2 + a ** 2 - a * lam_p) - (2 * r - 2) * (eta_p + (lam_p - a) ** 2)
    scale = r ** 4
    assert np.max(np.abs(R) / scale) < 1e-12 and np.max(np.abs(dR) / scale) < 1e-12
    # just outside the curve rays escape past a turning point, just inside they fall in
    sel = slice(20, -20, 40)
    for f, want in ((1 + 1e-6, 1), (1 - 1e-6, 0)):
        oc = O.is_outside_crit(a, th, al[sel] * f, be[sel] * f)
        assert np.all(oc == want), (f, oc)
