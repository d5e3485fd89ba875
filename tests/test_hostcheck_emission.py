"""The emission header (kg_emission.hpp) built for the CPU against the
oracle: special functions, every velocity / field / emissivity model, and
the polarized image for eleven model configurations (both observer
hemispheres, negative spin, cuts, pathlength, tables).  The GPU runs the
same source (tests/test_gpu_emission.py)."""

import mpmath as mp
import numpy as np
import pytest
from scipy import special

from kgeo_amd import models as M
from oracle import kgeo_oracle_emission as E

from . import emission_cases as C
from .kgeo_testlib import Host


@pytest.fixture(scope="module")
def host():
    return Host()


def test_special_functions(host):
    x = np.concatenate([-np.logspace(3, -6, 25), np.linspace(-1, 1, 41), [0.4999, 0.5, 0.5001, 1.0]])
    li, _ = host.special(0, x)
    ref = np.array([float(mp.polylog(2, v)) for v in x])
    assert np.max(np.abs(li - ref) / np.maximum(np.abs(ref), 1e-300)) < 4e-15
    w = np.concatenate([-1 / np.e + np.logspace(-10, -1, 10), np.linspace(-0.3, 5, 50), np.logspace(1, 12, 12)])
    lw, _ = host.special(1, w)
    ref = np.real(special.lambertw(w))
    cond = 1 + 1 / np.abs(1 + ref)   # relative condition number of W, 1/|1+W| (large at the branch point)
    assert np.all(np.abs(lw - ref) / np.maximum(np.abs(ref), 1e-3) < 1e-15 * cond + 1e-15)
    r = np.array([1.05, 1.3, 1.9999, 2.0, 2.0001, 2.6, 3.2, 3.9999, 4.0, 4.5, 7., 15., 40., 99.9, 100., 100.1, 300.])
    f, fp = host.special(2, r)
    assert np.allclose(f, E.bz_f(r), rtol=3e-14, atol=0)
    assert np.allclose(fp, E.bz_fprime(r), rtol=1e-12, atol=0)


def _models(vel, bf, em):
    return M.emission_model(vel, bf, em, device=False)


VEL = [("zamo", {}), ("infall", {}), ("kep", {}), ("kep", dict(retrograde=True)), ("subkep", dict(fac_subkep=0.7)),
       ("general", dict(fac_subkep=0.5, beta_r=0.3, beta_phi=0.6)), ("general", dict(retrograde=True)),
       ("gelles", dict(gelles_beta=0.4)), ("simfit", {}), ("fromfile", dict(file=C.TABLE)),
       ("driftframe", dict(nu_parallel=-0.2)), ("driftframe", dict(bfield=("bz_monopole", {}), nu_parallel=0.5))]
BF = [("rad", {}), ("vert", {}), ("tor", {}), ("simple", dict(Cr=0.3, Cvert=-1., Cph=2.)),
      ("simple_rm1", dict(Cr=1., Cph=-0.4)), ("monopoleA", {}), ("bz_monopole", {}),
      ("bz_monopole", dict(secondorder_only=True, C=-1.)), ("bz_guess", dict(C=-1.)), ("bz_para", {}),
      ("power", dict(p=0.6)), ("power", dict(p=0., usemono=True)), ("gen_power", dict(n_I=1., p_val=0.5)),
      ("fromfile", dict(file=C.TABLE))]


@pytest.mark.parametrize("vt,vkw", VEL, ids=lambda v: str(v))
def test_velocities_and_tetrads(host, vt, vkw):
    def mk(mod):
        kw = dict(vkw)
        if "bfield" in kw:
            kw["bfield"] = mod.Bfield(kw["bfield"][0], **kw["bfield"][1])
        return mod.Velocity(vt, **kw)
    ov, pv = mk(E), mk(M)
    for a in (0.3, -0.5, 0.9):
        rh = 1 + np.sqrt(1 - a * a)
        r = np.linspace(rh + 0.02, 25, 97)
        got = host.model_eval(_models(pv, M.Bfield("rad"), M.Emissivity("constant")), a, r)
        ref = np.array([np.broadcast_to(np.asarray(x, float), r.shape) for x in ov.u_lab(a, r)])
        ok = np.isfinite(ref).all(0)
        assert np.array_equal(np.isfinite(got["u"]).all(0), ok), (vt, a)
        case = ("", a, 0, 0, (vt, vkw), None, None, False, False)
        tol = 1e-11 + C.plunge_slack(case, r)     # the plunge formula's own conditioning
        err = np.abs(got["u"][:, ok] - ref[:, ok]) / np.maximum(np.abs(ref[:, ok]), 1e-3 * np.abs(ref[0, ok]))
        assert np.all(err <= tol[ok]), (vt, a, float(err.max()))
        tet = ov.tetrades(a, r)
        T = np.array([[np.broadcast_to(np.asarray(c, float), r.shape) for c in e] for e in tet]).reshape(16, -1)
        e2 = np.abs(got["tetrad"][:, ok] - T[:, ok]) / np.maximum(np.abs(T[:, ok]), 1e-3 * np.abs(T[:, ok]).max(0))
        assert np.all(e2 <= 1e3 * tol[ok]), (vt, a, float(e2.max()))


@pytest.mark.parametrize("bt,bkw", BF, ids=lambda v: str(v))
def test_fields_omega_and_bsq(host, bt, bkw):
    ob, pb = E.Bfield(bt, **bkw), M.Bfield(bt, **bkw)
    for a in (0.4, -0.8):
        rh = 1 + np.sqrt(1 - a * a)
        r = np.concatenate([np.linspace(rh + 0.01, 12, 60), [50., 99., 101., 200.]])
        for th in (None, np.full(r.shape, 1.1)):
            if th is not None and bt in ("rad", "vert", "tor", "simple", "simple_rm1", "fromfile"):
                continue   # equatorial-only models (bfields.py:437, 462)
            got = host.model_eval(_models(M.Velocity("zamo"), pb, M.Emissivity("constant")), a, r, th=th)
            ref = np.array([np.broadcast_to(np.asarray(x, float), r.shape)
                            for x in ob.bfield_lab(a, r, th=np.pi / 2 if th is None else 1.1)])
            scale = np.abs(ref).max(0)
            assert np.allclose(got["B"], ref, rtol=1e-12, atol=1e-13 * scale.max()), (bt, a)
            if bt in ("monopoleA", "bz_monopole", "bz_guess", "bz_para", "power", "gen_power"):
                om = np.broadcast_to(np.asarray(ob.omega_field(a, r, th=np.pi / 2 if th is None else 1.1), float),
                                     r.shape)
                assert np.allclose(got["omega"], om, rtol=1e-11, atol=1e-16), (bt, a)
        got = host.model_eval(_models(M.Velocity("kep"), pb, M.Emissivity("constant")), a, r)
        ref = ob.bsq(a, r, E.Velocity("kep"))
        assert np.allclose(got["bsq"], ref, rtol=1e-9, atol=1e-14 * np.abs(ref).max()), (bt, a)


@pytest.mark.parametrize("et,ekw", [("constant", {}), ("bpl", {}), ("bpl", dict(p1=-1.5, p2=-0.2)), ("glm", {}),
                                    ("ring", dict(r_ring=3.5, gamma_off=0.3, sigma=0.5)),
                                    ("thermal", C.THERMAL), ("thermal", dict(C.THERMAL, bfield_model=True))])
def test_emissivities(host, et, ekw):
    oe, pe = E.Emissivity(et, **ekw), M.Emissivity(et, **ekw)
    rng = np.random.default_rng(1)
    r = np.linspace(1.5, 30, 50)
    g = rng.uniform(0.3, 1.8, r.size)
    s = rng.uniform(0.05, 1, r.size)
    bm = rng.uniform(1, 50, r.size)
    for spec in (0., 1., 2.3):
        model = M.emission_model(M.Velocity("zamo"), M.Bfield("rad"), pe, specind=spec, nu_obs=86e9, device=False)
        got = host.model_eval(model, 0.6, r, g=g, sinthb=s, bmag=bm)["j"]
        ref = oe.jrest(0.6, r, g, s, Bmag=bm, nu_obs=86e9, specind=spec)
        assert np.allclose(got, ref, rtol=1e-12, atol=0), (et, spec)
        if et != "thermal":   # g and sinthetab None
            model = M.emission_model(M.Velocity("zamo"), M.Bfield("rad"), pe, specind=spec, device=False)
            got = host.model_eval(model, 0.6, r)["j"]
            assert np.allclose(got, oe.jrest(0.6, r), rtol=1e-12, atol=0), et


@pytest.mark.parametrize("case", C.CASES, ids=lambda c: c[0])
def test_polarized_image(host, case):
    (ov, ob, oe), (pv, pb, pe) = C.build(case, E, M)
    name, a, inc, mbar_max, _, _, _, pol, pl = case
    A, B = C.pixels(16)
    ref = C.oracle_image(case, ov, ob, oe, A, B)
    model = M.emission_model(pv, pb, pe, pol, pl, 1., 0.1, 230e9, device=False)
    got = host.stokes(a, np.inf, inc * np.pi / 180, model, A, B, mbar_max + 1)
    w = C.compare(case, got, ref)
    emitting = np.count_nonzero(np.asarray(ref["g"]) != 0)
    assert emitting > 40, (name, emitting)
    print(f"\n[{name}] {emitting} emitting samples; worst " + ", ".join(f"{k} {v:.1e}" for k, v in w.items()))
