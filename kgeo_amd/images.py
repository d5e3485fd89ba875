"""Polarized images of an equatorial source on the GPU: drop-ins for
make_image and Iobs (kgeo/equatorial_images.py:23-264).

One kgeo_equatorial_stokes launch computes every subring of every pixel: the
emission radius and Mino time of the mbar-th equatorial crossing
(r_equatorial), the masks, the fluid velocity and redshift, and with
polarization the comoving-frame field, sin(theta_B), the Penrose-Walker
constant and the EVPA; then the emissivity and I, Q, U (kg_emission.hpp).
The models are kgeo_amd.models objects, built with the reference's arguments.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .models import Bfield, Emissivity, Velocity, emission_model
from .raytrace import _as_device_f64, _device, _ptr, _stream_handle

bfield_default = Bfield("rad")          # equatorial_images.py:15-17
vel_default = Velocity("zamo")
emis_default = Emissivity("bpl")
SPECIND = 1        # equatorial_images.py:18-21
DISKANGLE = 0.1
SINTHB = np.pi / 4.
OBSFREQ = 230.e9


def _check(a, r_o, th_o, mbar, name="mbar_max"):
    """The argument checks of make_image / Iobs (equatorial_images.py:36-44,
    :144-152), with the reference's messages."""
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")
    if not (isinstance(r_o, float) and (r_o >= 100)):
        raise Exception("r_o should be a float > 100")
    if not (isinstance(th_o, float) and (0 < th_o < np.pi) and th_o != 0.5 * np.pi):
        raise Exception("th_o should be a float in range (0,pi/2) or (pi/2,pi)")
    if not (isinstance(mbar, int) and (mbar >= 0)):
        raise Exception("%s should be an integer >=0!" % name)


def _check_models(bfield, velocity, pathlength):
    if bfield.fieldframe != "lab":
        raise Exception("bfield.fluidframe=%s: comoving fields do not reach the image" % bfield.fieldframe)
    if pathlength and velocity.veltype == "fromfile" and np.any(velocity._table.host[2] != 0):
        raise Exception("calc_pathlength assumes that u2==0!")


def pixel_grid(alpha_min, alpha_max, beta_min, beta_max, psize):
    """make_image's screen grid (equatorial_images.py:49-57): n = int(floor(
    max - min) / psize) points from min to min + n*psize."""
    n_alpha = int(np.floor(alpha_max - alpha_min) / psize)
    alphas = np.linspace(alpha_min, alpha_min + n_alpha * psize, n_alpha)
    n_beta = int(np.floor(beta_max - beta_min) / psize)
    betas = np.linspace(beta_min, beta_min + n_beta * psize, n_beta)
    A, B = np.meshgrid(alphas, betas)
    return A.flatten(), B.flatten()


def stokes_device(a, r_o, th_o, mbar_max, alpha, beta, emissivity=emis_default, velocity=vel_default,
                  bfield=bfield_default, polarization=False, pathlength=False, specind=SPECIND, diskangle=DISKANGLE,
                  nu_obs=OBSFREQ, want=("I", "Q", "U", "g", "r_s", "sin_thb", "Ir", "nmax_eq", "n_poloidal")):
    """kgeo_equatorial_stokes on device arrays: dict of GPU tensors, per-
    subring outputs (npix, mbar_max + 1), nmax_eq and n_poloidal (npix,).
    Asynchronous on the current stream."""
    lib = _lib.load()
    dev = _device()
    _check_models(bfield, velocity, pathlength)
    al = _as_device_f64(alpha, dev).reshape(-1)
    be = _as_device_f64(beta, dev).reshape(-1)
    n = al.numel()
    ns = int(mbar_max) + 1
    f64 = dict(device=dev, dtype=torch.float64)
    per_pixel = ("nmax_eq", "n_poloidal")
    out = {k: torch.empty((n,) if k in per_pixel else (n, ns), **f64) for k in want}
    model = emission_model(velocity, bfield, emissivity, polarization, pathlength, specind, diskangle, nu_obs)
    prm = _lib.make_params(float(a), float(r_o), float(th_o), n, 1, 1.0)
    names = ("I", "Q", "U", "g", "r_s", "sin_thb", "Ir", "nmax_eq", "n_poloidal", "lp", "Ie")
    rc = lib.kgeo_equatorial_stokes(prm, _ptr(al), _ptr(be), int(mbar_max), ctypes.byref(model),
                                    *[_ptr(out.get(k)) for k in names], _stream_handle())
    _lib.check(rc, "kgeo_equatorial_stokes")
    out["_keep"] = (velocity, bfield)   # tables stay alive until the launch has run
    return out


def _host(out, keys):
    torch.cuda.current_stream().synchronize()
    return [out[k].cpu().numpy() for k in keys]


def make_image(a, r_o, th_o, mbar_max, alpha_min, alpha_max, beta_min, beta_max, psize, nmax_only=False,
               emissivity=emis_default, bfield=bfield_default, velocity=vel_default, polarization=False,
               pathlength=False, specind=SPECIND, diskangle=DISKANGLE, nu_obs=OBSFREQ):
    """Drop-in for make_image (equatorial_images.py:23-122): (I, Q, U, r, t,
    g, sinthb, n, np[, lp, Ie]) with per-subring arrays (npix, mbar_max + 1)
    over the reference's pixel grid.  nmax_only=True returns zeros but n and
    np (the reference's own nmax_only branch returns an unassigned name)."""
    _check(a, r_o, th_o, mbar_max)
    alpha, beta = pixel_grid(alpha_min, alpha_max, beta_min, beta_max, psize)
    return make_image_pixels(a, r_o, th_o, mbar_max, alpha, beta, nmax_only=nmax_only, emissivity=emissivity,
                             bfield=bfield, velocity=velocity, polarization=polarization, pathlength=pathlength,
                             specind=specind, diskangle=diskangle, nu_obs=nu_obs)


def make_image_pixels(a, r_o, th_o, mbar_max, alpha, beta, nmax_only=False, emissivity=emis_default,
                      bfield=bfield_default, velocity=vel_default, polarization=False, pathlength=False,
                      specind=SPECIND, diskangle=DISKANGLE, nu_obs=OBSFREQ):
    """make_image on caller-given screen points."""
    _check(a, r_o, th_o, mbar_max)
    n = len(np.asarray(alpha).reshape(-1))
    ns = mbar_max + 1
    if nmax_only:
        out = stokes_device(a, r_o, th_o, mbar_max, alpha, beta, emissivity, velocity, bfield, polarization,
                            pathlength, specind, diskangle, nu_obs, want=("nmax_eq", "n_poloidal"))
        nmax, npol = _host(out, ("nmax_eq", "n_poloidal"))
        z = [np.zeros((n, ns)) for _ in range(7)]
        return tuple(z) + (nmax, npol)
    keys = ["I", "Q", "U", "r_s", "Ir", "g", "sin_thb", "nmax_eq", "n_poloidal"]
    if pathlength:
        keys += ["lp", "Ie"]
    out = stokes_device(a, r_o, th_o, mbar_max, alpha, beta, emissivity, velocity, bfield, polarization, pathlength,
                        specind, diskangle, nu_obs, want=tuple(keys))
    return tuple(_host(out, keys))


def Iobs(a, r_o, th_o, mbar, alpha, beta, emissivity=emis_default, velocity=vel_default, bfield=bfield_default,
         polarization=False, pathlength=False, specind=SPECIND, diskangle=DISKANGLE, nu_obs=OBSFREQ,
         th_s=np.pi / 2):
    """Drop-in for Iobs (equatorial_images.py:124-264): (I, Q, U, g, r_s,
    sin_thb, Ir, Imax, Nmax[, lp, Ie]) of subring mbar."""
    _check(a, r_o, th_o, mbar, name="mbar")
    if th_s != np.pi / 2:
        raise NotImplementedError("Iobs evaluates the emission at th_s = pi/2 only")
    alpha = alpha if isinstance(alpha, np.ndarray) else np.array([alpha]).flatten()
    beta = beta if isinstance(beta, np.ndarray) else np.array([beta]).flatten()
    if len(alpha) != len(beta):
        raise Exception("alpha, beta are different lengths!")
    keys = ["I", "Q", "U", "g", "r_s", "sin_thb", "Ir", "nmax_eq"] + (["lp", "Ie"] if pathlength else [])
    out = stokes_device(a, r_o, th_o, mbar, alpha, beta, emissivity, velocity, bfield, polarization, pathlength,
                        specind, diskangle, nu_obs, want=tuple(keys))
    # I_max: the total Mino time of r_equatorial, from the equatorial image kernel
    from .equatorial import equatorial_image_device
    imax = equatorial_image_device(a, r_o, th_o, alpha, beta, mbar_max=0)["Imax"]
    vals = dict(zip(keys, _host(out, keys)))
    torch.cuda.current_stream().synchronize()
    Imax = imax.cpu().numpy()
    res = [vals[k][:, mbar] for k in ("I", "Q", "U", "g", "r_s", "sin_thb", "Ir")] + [Imax, vals["nmax_eq"]]
    if pathlength:
        res += [vals["lp"][:, mbar], vals["Ie"][:, mbar]]
    return tuple(res)
