"""Lensed curves of constant equatorial radius on the GPU
(equatorial_lensing.py:198-423).

The reference finds, for each screen angle varphi, the screen radius rho
at which the mbar-th equatorial crossing of the ray lands on r = req, by a
scipy brentq per angle (rho_of_req, :226-262), and builds a library of such
curves over a spin x inclination grid with a 10-process Pool
(generate_library, :351-423).  Here every (curve, angle) pair is one lane of
one kernel (kgeo_rho_of_req): Brent's method over the reference's bracket
with brentq's tolerances, the objective evaluated with the device
r_equatorial.  critical_curve (:264-293) evaluates its closed form per
photon-orbit radius on the GPU (kgeo_critical_curve).
"""

import ctypes
import os

import numpy as np
import torch

from . import _lib
from .raytrace import _device, _ptr, _stream_handle

INF = 1.e50           # equatorial_lensing.py:18
R0 = float("inf")     # equatorial_lensing.py:19: the observer at infinity

ST_FOUND, ST_NO_BRACKET, ST_NOT_CONVERGED = 0, 1, 2


def default_varphis():
    """The reference's 360 screen angles (equatorial_lensing.py:227-228)."""
    return np.linspace(-180, 179, 360) * np.pi / 180.


def _curve_params(spins, incs):
    """A device array of kgeo_ray_params, one per (spin, inclination) curve,
    with numpy's derived scalars and the observer at infinity."""
    n = len(spins)
    arr = (_lib.RayParams * n)()
    for k in range(n):
        arr[k] = _lib.make_params(float(spins[k]), R0, float(incs[k]), 1, 1, 1.0)
    raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
    return torch.from_numpy(raw).to(_device())


def rho_of_req_device(spins, incs, reqs, mbars, varphis):
    """Screen radii of many curves at once: curve k is the image of the
    equatorial circle r = reqs[k] in subring mbars[k] for spin spins[k] and
    inclination incs[k] (radians).  Returns (rho, status), tensors of shape
    (ncurves, len(varphis)); status 0 where a root was found, 1 where the
    bracket holds no sign change (brentq raises ValueError), 2 where Brent's
    method did not converge in 100 evaluations (RuntimeError)."""
    lib = _lib.load()
    dev = _device()
    spins = np.atleast_1d(np.asarray(spins, dtype=np.float64))
    n = len(spins)
    incs = np.broadcast_to(np.asarray(incs, dtype=np.float64), (n,))
    reqs = torch.from_numpy(np.array(np.broadcast_to(np.asarray(reqs, np.float64), (n,)))).to(dev)
    mbars = torch.from_numpy(np.array(np.broadcast_to(np.asarray(mbars, np.int32), (n,)))).to(dev)
    vp = torch.from_numpy(np.ascontiguousarray(varphis, dtype=np.float64)).to(dev)
    curves = _curve_params(spins, incs)
    rho = torch.empty((n, vp.numel()), device=dev, dtype=torch.float64)
    status = torch.empty((n, vp.numel()), device=dev, dtype=torch.int32)
    _lib.check(lib.kgeo_rho_of_req(_ptr(curves), _ptr(reqs), _ptr(mbars), n, _ptr(vp), vp.numel(), _ptr(rho),
                                   _ptr(status), _stream_handle()), "kgeo_rho_of_req")
    return rho, status


def _raise_for(status):
    if np.any(status == ST_NO_BRACKET):
        raise ValueError("f(a) and f(b) must have different signs")
    if np.any(status == ST_NOT_CONVERGED):
        raise RuntimeError("Failed to converge after 100 iterations")


def _check(a, th0):
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")
    if not (isinstance(th0, float) and (0 < th0 < np.pi) and th0 != 0.5 * np.pi):
        raise Exception("th_o should be a float in range (0,pi/2) or (pi/2,pi)")


def rho_of_req(a, th0, req, mbar=0, varphis=None):
    """Drop-in for equatorial_lensing.rho_of_req (:226-247): (varphis, rhos,
    alphas, betas) of the image of r = req in subring mbar.  Raises as brentq
    does where an angle has no root in the bracket."""
    _check(a, th0)
    if varphis is None:
        varphis = default_varphis()
    varphis = np.asarray(varphis, dtype=np.float64)
    rho, status = rho_of_req_device([a], [th0], [float(req)], [int(mbar)], varphis)
    torch.cuda.current_stream().synchronize()
    rhos = rho[0].cpu().numpy()
    _raise_for(status[0].cpu().numpy())
    return varphis, rhos, np.cos(varphis) * rhos, np.sin(varphis) * rhos


def rho_of_req_single(a, th0, req, varphi, mbar=0):
    """Drop-in for equatorial_lensing.rho_of_req_single (:249-262)."""
    _, rhos, _, _ = rho_of_req(a, th0, req, mbar, np.array([float(varphi)]))
    return float(rhos[0])


def critical_curve(a, th0, n=100000):
    """Drop-in for equatorial_lensing.critical_curve (:264-293): the
    critical curve's (alpha, beta) points, beta >= 0 half from the photon
    orbits r in [r_-, r_+] and its mirror image, NaN points dropped."""
    lib = _lib.load()
    dev = _device()
    rminus = 2 * (1 + np.cos(2. * np.arccos(-a) / 3.))
    rplus = 2 * (1 + np.cos(2. * np.arccos(a) / 3.))
    r = torch.from_numpy(np.linspace(rminus, rplus, n)).to(dev)
    al = torch.empty_like(r)
    be = torch.empty_like(r)
    p = _lib.make_params(float(a), R0, float(th0), 1, 1, 1.0)
    _lib.check(lib.kgeo_critical_curve(p, _ptr(r), n, _ptr(al), _ptr(be), _stream_handle()), "kgeo_critical_curve")
    torch.cuda.current_stream().synchronize()
    alpha, beta = al.cpu().numpy(), be.cpu().numpy()
    mask = np.isnan(alpha) + np.isnan(beta)
    alpha, beta = alpha[~mask], beta[~mask]
    return np.hstack((alpha, np.flip(alpha))), np.hstack((beta, -np.flip(beta)))


LIBRARY_RADII = {"r10": 10., "r100": 100., "r1000": 1000., "r10000": 10000.}   # :384-394


def generate_library(which="rh", outpath=None, spins=None, incs=None, mmax=5, varphis=None, max_lanes=1 << 24):
    """The reference's curve library (equatorial_lensing.py:351-423) in few
    launches: for every spin in spins (default linspace(0.01, 0.99, 99)) and
    inclination in incs (degrees, default 1..89), the curves mbar = 0..mmax
    of r = r_h (which='rh') or r = 10, 100, 1e3, 1e4, written as
    outpath/aSS_iII_<which>.txt with columns varphi, rho(mbar = 0..mmax).
    Existing files are kept; a (spin, inclination) with an angle that has no
    root writes no file, as the reference's try/except does.  Returns the
    list of files written."""
    if which != "rh" and which not in LIBRARY_RADII:
        raise Exception()
    outpath = outpath or ("./curve_library_" + which)
    os.makedirs(outpath, exist_ok=True)
    spins = np.linspace(0.01, 0.99, 99) if spins is None else np.asarray(spins, dtype=np.float64)
    incs = np.linspace(1, 89, 89) if incs is None else np.asarray(incs, dtype=np.float64)
    varphis = default_varphis() if varphis is None else np.asarray(varphis, dtype=np.float64)
    allincs, allspins = np.meshgrid(incs, spins)
    allincs, allspins = allincs.flatten(), allspins.flatten()

    def fname(spin, inc):
        return "%s/a%02.0f_i%02.0f_%s.txt" % (outpath, 100 * spin, inc, which)

    todo = [k for k in range(len(allspins)) if not os.path.isfile(fname(allspins[k], allincs[k]))]
    nm = mmax + 1
    per_launch = max(1, max_lanes // (nm * len(varphis)))
    written = []
    for s0 in range(0, len(todo), per_launch):
        idx = todo[s0:s0 + per_launch]
        sp = np.repeat(allspins[idx], nm)
        th = np.repeat(allincs[idx] * np.pi / 180., nm)
        rh = 1 + np.sqrt(1 - sp ** 2)
        req = rh if which == "rh" else np.full(len(sp), LIBRARY_RADII[which])
        mb = np.tile(np.arange(nm, dtype=np.int32), len(idx))
        rho, status = rho_of_req_device(sp, th, req, mb, varphis)
        torch.cuda.current_stream().synchronize()
        rho = rho.cpu().numpy().reshape(len(idx), nm, len(varphis))
        status = status.cpu().numpy().reshape(len(idx), nm, len(varphis))
        for q, k in enumerate(idx):
            if np.any(status[q] != ST_FOUND):
                continue
            out = np.hstack((varphis.reshape(-1, 1), rho[q].T))
            np.savetxt(fname(allspins[k], allincs[k]), out)
            written.append(fname(allspins[k], allincs[k]))
    return written
