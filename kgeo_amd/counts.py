"""Equatorial-crossing and poloidal-orbit counts on the GPU
(kerr_raytracing_utils.py:551-684)."""

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
