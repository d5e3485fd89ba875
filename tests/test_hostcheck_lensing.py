"""The lensed-curve kernels' source (kg_lensing.hpp: r_equatorial2, Brent's
method over the reference's bracket) on the CPU against the oracle
(equatorial_lensing.py:198-262)."""

import ctypes

import numpy as np
import pytest

from oracle import kgeo_oracle as O

from .kgeo_testlib import Host


def host_rho(h, a, th, req, mbar, vp):
    rho = np.empty(len(vp))
    st = np.empty(len(vp), np.int32)
    h.lib.kgh_rho_of_req(Host._p(Host.scalars(a, np.inf, th)), ctypes.c_double(req), ctypes.c_int(mbar),
                         Host._p(np.ascontiguousarray(vp)), ctypes.c_int64(len(vp)), Host._p(rho),
                         Host._p(st, ctypes.c_int32))
    return rho, st


def roots_agree(a, th, req, mbar, vp, rho, rho_ref, tol=2e-11):
    """Each root equals the oracle's to tol, or the oracle's own objective
    changes sign within 1e-9 of it (a root where the reference's objective is
    not resolved to 1e-10 itself, e.g. on beta = +-6e-16)."""
    bad = []
    for v, r, q in zip(vp, rho, rho_ref):
        if abs(r - q) <= tol * max(1.0, abs(q)):
            continue
        lo, mid, hi = (O.r_equatorial2(r + s, v, a, th, mbar) - req for s in (-1e-9, 0.0, 1e-9))
        if lo * mid > 0 and mid * hi > 0:
            bad.append((v, r, q))
    return bad


@pytest.mark.parametrize("a,th_deg,req,mbar", [(0.5, 30.0, 6.0, 0), (0.9, 60.0, None, 1), (0.3, 17.0, 10.0, 2),
                                               (-0.7, 80.0, 4.0, 0), (0.99, 85.0, None, 0), (0.0, 45.0, 5.0, 1)])
def test_rho_of_req_matches_oracle(a, th_deg, req, mbar):
    th = np.deg2rad(th_deg)
    req = (1 + np.sqrt(1 - a * a)) if req is None else req
    vp = np.linspace(-180, 179, 60) * np.pi / 180
    _, rho_ref, _, _ = O.rho_of_req(a, th, req, mbar, vp)
    rho, st = host_rho(Host(), a, th, req, mbar, vp)
    assert np.all(st == 0)
    assert not roots_agree(a, th, req, mbar, vp, rho, rho_ref)


def test_bracket_without_sign_change_is_reported():
    # r = 0.5 lies inside the horizon: r_equatorial2 - 0.5 > 0 over the whole
    # bracket (rays inside the critical curve map to 1), so brentq raises
    th = np.deg2rad(40.0)
    vp = np.array([0.3, 1.2])
    with pytest.raises(ValueError):
        O.rho_of_req(0.5, th, 0.5, 0, vp)
    rho, st = host_rho(Host(), 0.5, th, 0.5, 0, vp)
    assert np.all(st == 1) and np.all(np.isnan(rho))


def test_r_equatorial2_matches_oracle():
    rng = np.random.default_rng(4)
    h = Host()
    for a, thd, mbar in ((0.7, 35.0, 0), (-0.4, 70.0, 1), (0.95, 10.0, 2)):
        th = np.deg2rad(thd)
        rho = rng.uniform(0.0, 12.0, 300)
        vp = rng.uniform(-np.pi, np.pi, 300)
        got = np.empty(300)
        h.lib.kgh_r_equatorial2(Host._p(Host.scalars(a, np.inf, th)), ctypes.c_int(mbar),
                                Host._p(rho), Host._p(vp), ctypes.c_int64(300), Host._p(got))
        want = np.array([O.r_equatorial2(r, v, a, th, mbar) for r, v in zip(rho, vp)])
        assert np.array_equal(got == O.LENS_INF, want == O.LENS_INF)
        assert np.array_equal(got == 1.0, want == 1.0)
        assert np.max(np.abs(got - want) / np.abs(want)) < 1e-10
