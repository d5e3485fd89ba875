"""Equatorial-crossing, poloidal-orbit and turning-point counts on the GPU
(kerr_raytracing_utils.py:551-826)."""

import numpy as np
import torch

from . import _lib
from .raytrace import _as_device_f64, _device, _ptr, _stream_handle


def crossings_device(a, th_o, alpha, beta, tau, r_o=1.0e4):
    """(n_equatorial, n_poloidal) tensors of tau's shape ((ntau, npix) or (npix,))."""
    lib = _lib.load()
    dev = _device()
    al = _as_device_f64(alpha, dev).reshape(-1)
    be = _as_device_f64(beta, dev).reshape(-1)
    ta = _as_device_f64(tau, dev)
    one = ta.dim() == 1
    ta2 = ta.reshape(1, -1) if one else ta
    npix = al.numel()
    if ta2.shape[-1] != npix:
        raise Exception("tau has incompatible shape in n_poloidal_orbits!")
    neq = torch.empty_like(ta2)
    npol = torch.empty_like(ta2)
    p = _lib.make_params(a, r_o, th_o, npix, ta2.shape[0], 1.0)
    _lib.check(lib.kgeo_crossings(p, _ptr(al), _ptr(be), _ptr(ta2), ta2.shape[0], _ptr(neq), _ptr(npol),
                                  _stream_handle()), "kgeo_crossings")
    if one:
        return neq[0], npol[0]
    return neq, npol


def n_equatorial_crossings(a, th_o, alpha, beta, tau):
    """Floor-valued equatorial crossings (kerr_raytracing_utils.py:605-684), numpy out."""
    neq, _ = crossings_device(a, th_o, np.atleast_1d(alpha), np.atleast_1d(beta), np.asarray(tau, dtype=float))
    torch.cuda.current_stream().synchronize()
    return neq.cpu().numpy()


def n_poloidal_orbits(a, th_o, alpha, beta, tau):
    """Fractional poloidal orbits (kerr_raytracing_utils.py:551-603), numpy out."""
    _, npol = crossings_device(a, th_o, np.atleast_1d(alpha), np.atleast_1d(beta), np.asarray(tau, dtype=float))
    torch.cuda.current_stream().synchronize()
    return npol.cpu().numpy()


def turnings_device(a, th_o, alpha, beta, tau, r_o=1.0e4, want=("n_angular", "turns_theta", "turns_r")):
    """Turning-point counts at Mino times tau ((ntau, npix) or (npix,)), as
    tensors of tau's shape: n_angular (the reference's fractional
    n_angular_turnings, kerr_raytracing_utils.py:686-792), turns_theta (its
    floor, int32: theta turning points passed) and turns_r (int8: 1 once a
    ray outside the critical curve has passed its radial turning point at
    tau_tot / 2).  n_angular and turns_theta need th_o <= pi/2, as the
    reference does (:690-691)."""
    lib = _lib.load()
    dev = _device()
    al = _as_device_f64(alpha, dev).reshape(-1)
    be = _as_device_f64(beta, dev).reshape(-1)
    ta = _as_device_f64(tau, dev)
    one = ta.dim() == 1
    ta2 = ta.reshape(1, -1) if one else ta
    npix = al.numel()
    if ta2.shape[-1] != npix:
        raise Exception("tau has incompatible shape in n_poloidal_orbits!")
    out = {}
    if "n_angular" in want:
        out["n_angular"] = torch.empty_like(ta2)
    if "turns_theta" in want:
        out["turns_theta"] = torch.empty(ta2.shape, device=dev, dtype=torch.int32)
    if "turns_r" in want:
        out["turns_r"] = torch.empty(ta2.shape, device=dev, dtype=torch.int8)
    p = _lib.make_params(a, r_o, th_o, npix, ta2.shape[0], 1.0)
    _lib.check(lib.kgeo_turnings(p, _ptr(al), _ptr(be), _ptr(ta2), ta2.shape[0], _ptr(out.get("n_angular")),
                                 _ptr(out.get("turns_theta")), _ptr(out.get("turns_r")), _stream_handle()),
               "kgeo_turnings")
    if one:
        out = {k: v[0] for k, v in out.items()}
    return out


def n_angular_turnings(a, th_o, alpha, beta, tau):
    """Fractional angular turnings (kerr_raytracing_utils.py:686-792), numpy out;
    the reference's argument checks and messages (:689-699)."""
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")
    if not (isinstance(th_o, float) and (0 <= th_o <= np.pi / 2.)):
        raise Exception("th_o should be a float in range (0,pi/2]")
    alpha, beta = np.atleast_1d(alpha), np.atleast_1d(beta)
    if len(alpha) != len(beta):
        raise Exception("alpha, beta are different lengths!")
    out = turnings_device(a, th_o, alpha, beta, np.asarray(tau, dtype=float), want=("n_angular",))
    torch.cuda.current_stream().synchronize()
    return out["n_angular"].cpu().numpy()


def ellip_pi_device(n, phi, m):
    """Pi(n; phi | m) in the A&S convention with the special cases of the
    reference's ellippi_arr plug-in (scipy_ellip_binding.py:87-237).  n and m
    must already have phi's shape (the reference broadcasts them with
    np.outer, :102-106).  Returns a tensor of phi's shape."""
    lib = _lib.load()
    dev = _device()
    ph = _as_device_f64(phi, dev)
    nn = _as_device_f64(n, dev)
    mm = _as_device_f64(m, dev)
    if nn.shape != ph.shape or mm.shape != ph.shape:
        raise ValueError("n, phi and m must have the same shape")
    out = torch.empty_like(ph)
    _lib.check(lib.kgeo_ellip_pi(_ptr(nn), _ptr(ph), _ptr(mm), _ptr(out), ph.numel(), _stream_handle()),
               "kgeo_ellip_pi")
    return out


def ellippi_arr(n, phi, m):
    """Drop-in for the reference's ellip_pi_arr(n, phi, m), numpy in and out
    (scipy_ellip_binding.py:87-237): scalars become 1-element arrays, the
    three last dimensions must agree (same message as :99-100), and 1-D n or
    m is repeated along phi's leading axis (:103-106)."""
    n, phi, m = (np.asarray(x, dtype=np.float64).reshape(-1) if np.ndim(x) == 0 else
                 np.asarray(x, dtype=np.float64) for x in (n, phi, m))
    if not (n.shape[-1] == phi.shape[-1] == m.shape[-1]):
        raise Exception("inputs to ellip_pi_arr do not have same last dimension!")
    if n.ndim != phi.ndim:
        n = np.outer(np.ones(phi.shape[0]), n)
    if m.ndim != phi.ndim:
        m = np.outer(np.ones(phi.shape[0]), m)
    out = ellip_pi_device(np.ascontiguousarray(n), np.ascontiguousarray(phi), np.ascontiguousarray(m))
    torch.cuda.current_stream().synchronize()
    return out.cpu().numpy()
