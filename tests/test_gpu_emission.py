"""The polarized equatorial image and the emission models on the GPU,
through the C-ABI (kgeo_equatorial_stokes, kgeo_model_eval):

* eleven model configurations against the oracle's make_image (every output,
  1e-10 relative plus the plunge-formula slack inside the ISCO; N_max exact);
* the reference's own known answers (emissivities, bz_guess, bz_monopole,
  fromfile, general velocity) through the drop-in model classes, and
  u.u = -1 for every velocity model;
* the reference's tests/test_images.py in full: 800 x 800 images of subrings
  0 and 1 by make_image, EVPA cuts against ipole;
* make_image / Iobs drop-in semantics and refusals."""

import numpy as np
import pytest

from oracle import kgeo_oracle_emission as E

from . import emission_cases as C
from . import test_oracle_emission as P

pytestmark = pytest.mark.gpu


def _host(out, keys):
    import torch
    torch.cuda.synchronize()
    return {k: out[k].cpu().numpy() for k in keys}


@pytest.mark.parametrize("case", C.CASES, ids=lambda c: c[0])
def test_gpu_polarized_image(case):
    from kgeo_amd import images
    from kgeo_amd import models as M
    (ov, ob, oe), (pv, pb, pe) = C.build(case, E, M)
    name, a, inc, mbar_max, _, _, _, pol, pl = case
    A, B = C.pixels(24, seed=5)
    ref = C.oracle_image(case, ov, ob, oe, A, B)
    print(f"\n[{name}] {len(A)} px, subrings 0..{mbar_max}", flush=True)
    out = images.stokes_device(a, np.inf, inc * np.pi / 180, mbar_max, A, B, emissivity=pe, velocity=pv, bfield=pb,
                               polarization=pol, pathlength=pl, specind=1., want=tuple(ref))
    got = _host(out, ref)
    w = C.compare(case, got, ref)
    print(", ".join(f"{k} {v:.1e}" for k, v in w.items()))


def test_gpu_model_known_answers():
    from kgeo_amd import models as M
    for model, spin, r, kw, answer in P.EMIS_CASES:
        assert np.isclose(M.Emissivity(model, **kw).jrest(float(spin), np.array(r)), answer, rtol=1e-4)
    for Cc, spin, r, ratio in P.BZ_GUESS_CASES:
        B = np.squeeze(M.Bfield("bz_guess", C=Cc).bfield_lab(spin, r, th=np.pi / 2))
        assert np.isclose(B[2] / B[0], ratio, rtol=1e-5)
    for Cc, spin, r, ref in P.BZ_MONO_CASES:
        B = np.squeeze(M.Bfield("bz_monopole", C=Cc).bfield_lab(spin, r, th=np.pi / 2))
        assert np.allclose(B, ref, rtol=1e-3, atol=1e-15)
    b = M.Bfield("fromfile", file=C.TABLE)
    for r, ref in P.FROMFILE_B:
        assert np.allclose(np.squeeze(b.bfield_lab(0.6, r)), ref, rtol=1e-4)
    v = M.Velocity("fromfile", file=C.TABLE)
    for r, ref in P.FROMFILE_U:
        assert np.allclose(np.squeeze(v.u_lab(0.6, r)), ref, rtol=1e-4)
    for (retro, fk, bphi, br, spin, r), ref in P.GENERAL_CASES:
        v = M.Velocity("general", retrograde=bool(retro), fac_subkep=fk, beta_phi=bphi, beta_r=br)
        assert np.allclose(np.squeeze(v.u_lab(spin, r)), ref, rtol=1e-5)


def test_gpu_velocity_normalisation():
    from kgeo_amd import models as M
    rng = np.random.default_rng(42)
    for veltype in P.VELOCITY_MODELS:
        v = M.Velocity("driftframe", bfield=M.Bfield("bz_monopole"), nu_parallel=0.3) if veltype == "driftframe" \
            else M.Velocity(veltype)
        for spin in rng.random(3):
            reh = 1 + np.sqrt(1 - spin ** 2)
            r = np.linspace(reh + 0.05, 15., 40)
            u = np.array(v.u_lab(float(spin), r))
            for j in range(len(r)):
                if np.all(np.isfinite(u[:, j])):
                    g = P.gcov(spin, r[j])
                    assert np.isclose(u[:, j] @ g @ u[:, j], -1., rtol=1e-9), (veltype, spin, r[j])


@pytest.mark.parametrize("case", P.IPOLE, ids=lambda c: c[0])
def test_gpu_evpa_against_ipole(case):
    """The reference's tests/test_images.py, run through make_image at its
    full 800 x 800 size."""
    from kgeo_amd import images
    from kgeo_amd import models as M

    def image(a, th_o, vkw, al, be):
        out = images.make_image(a, np.inf, th_o, 1, -10., 10., -10., 10., 20. / P.NX, nmax_only=False,
                                emissivity=M.Emissivity("ring"), velocity=M.Velocity("general", **vkw),
                                bfield=M.Bfield("bz_guess", C=1.), polarization=True, specind=1.)
        Q, U = out[1], out[2]
        assert Q.shape == (P.NX * P.NX, 2)
        idx = np.concatenate([400 * P.NX + np.arange(P.NX), np.arange(P.NX) * P.NX + 400])
        # the cut pixels are the ones the caller asked for
        A, B = images.pixel_grid(-10., 10., -10., 10., 20. / P.NX)
        assert np.array_equal(A[idx], al) and np.array_equal(B[idx], be)
        return Q[idx], U[idx]

    P.ipole_cuts(case, image)


def test_gpu_make_image_and_iobs_dropins():
    from kgeo_amd import _lib, images
    from kgeo_amd import models as M
    a, th_o = 0.5, 1.1
    out = images.make_image(a, np.inf, th_o, 2, -6., 6., -5., 5., 0.5, emissivity=M.Emissivity("glm"),
                            velocity=M.Velocity("kep"), bfield=M.Bfield("bz_guess"), polarization=True,
                            pathlength=True)
    assert len(out) == 11
    A, B = E.pixel_grid(-6., 6., -5., 5., 0.5)
    ref = E.make_image_pixels(a, np.inf, th_o, 2, A, B, emissivity=E.Emissivity("glm"), velocity=E.Velocity("kep"),
                              bfield=E.Bfield("bz_guess"), polarization=True, pathlength=True, specind=1)
    case = ("", a, 0, 2, ("kep", {}), None, None, True, True)
    keys = ["I", "Q", "U", "r_s", "Ir", "g", "sin_thb", "nmax_eq", "n_poloidal", "lp", "Ie"]
    C.compare(case, dict(zip(keys, out)), dict(zip(keys, ref)))
    nm = images.make_image(a, np.inf, th_o, 1, -6., 6., -5., 5., 0.5, nmax_only=True)
    assert np.array_equal(nm[7], ref[7]) and np.allclose(nm[8], ref[8], rtol=1e-12) and not np.any(nm[0])
    d = images.Iobs(a, np.inf, th_o, 1, A, B, emissivity=M.Emissivity("glm"), velocity=M.Velocity("kep"),
                    bfield=M.Bfield("bz_guess"), polarization=True)
    r = E.Iobs(a, np.inf, th_o, 1, A, B, E.Emissivity("glm"), E.Velocity("kep"), E.Bfield("bz_guess"),
               polarization=True)
    assert len(d) == 9
    for k, (x, y) in enumerate(zip(d, r)):
        ok = np.isfinite(y)
        assert np.array_equal(np.isnan(x), np.isnan(y)), k
        assert np.allclose(x[ok], y[ok], rtol=1e-10, atol=1e-10 * np.abs(y[ok]).max(initial=0)), k
    # refusals, with the reference's messages where it has one
    with pytest.raises(Exception, match="r_o should be a float > 100"):
        images.make_image(a, 50., th_o, 1, -6., 6., -5., 5., 0.5)
    with pytest.raises(Exception, match="th_o should be"):
        images.make_image(a, np.inf, np.pi / 2, 1, -6., 6., -5., 5., 0.5)
    with pytest.raises(_lib.KgeoError, match="u2==0"):
        images.make_image(a, np.inf, th_o, 1, -6., 6., -5., 5., 0.5, pathlength=True,
                          velocity=M.Velocity("driftframe", bfield=M.Bfield("bz_monopole")))
    with pytest.raises(Exception, match="comoving"):
        images.make_image(a, np.inf, th_o, 1, -6., 6., -5., 5., 0.5, bfield=M.Bfield("const_comoving", Cr=1.))
    with pytest.raises(_lib.KgeoError, match="fac_subkep"):
        images.make_image(a, np.inf, th_o, 1, -6., 6., -5., 5., 0.5, velocity=M.Velocity("subkep", fac_subkep=1.5))
