"""Equatorial-emission image and its gradient on the MI355X kernels.

image(alpha, beta) = sum_{mbar=0..mbar_max} g^3 r_s^-p over the subrings that
emit, with r_s from the mbar-th equatorial crossing (equatorial_lensing.py
r_equatorial :85-196), g the redshift of a Keplerian disk with Cunningham
plunge inside the ISCO (velocities.py u_kep :229-281, equatorial_images.py
calc_redshift :307-334) and the reference's emission masks
(equatorial_images.py :175, :266-305).

EquatorialImageFn is a torch.autograd.Function over the parameter vector
(a, th_o, p); its backward is the tangent kernel (forward-mode duals, the
geodesic recomputed per pixel).
"""

import numpy as np
import torch

from . import _lib
from .raytrace import _as_device_f64, _device, _ptr, _stream_handle

MBAR_MAX = 2


def _params(a, r_o, th_o, npix):
    return _lib.make_params(float(a), float(r_o), float(th_o), npix, 1, 1.0)


def equatorial_image_device(a, r_o, th_o, alpha, beta, mbar_max=MBAR_MAX, p=2.0, want_parts=True):
    """Image and per-subring outputs as GPU tensors (async on the current stream)."""
    lib = _lib.load()
    dev = _device()
    al = _as_device_f64(alpha, dev).reshape(-1)
    be = _as_device_f64(beta, dev).reshape(-1)
    n = al.numel()
    M = mbar_max + 1
    f64 = dict(device=dev, dtype=torch.float64)
    out = dict(image=torch.empty(n, **f64))
    if want_parts:
        out.update(I=torch.empty((M, n), **f64), g=torch.empty((M, n), **f64), r_s=torch.empty((M, n), **f64),
                   Ir=torch.empty((M, n), **f64), Imax=torch.empty(n, **f64),
                   Nmax=torch.empty(n, device=dev, dtype=torch.int32),
                   zero_mask=torch.empty(n, device=dev, dtype=torch.uint8))
    prm = _params(a, r_o, th_o, n)
    rc = lib.kgeo_equatorial_image(prm, _ptr(al), _ptr(be), int(mbar_max), float(p), _ptr(out["image"]),
                                   _ptr(out.get("I")), _ptr(out.get("g")), _ptr(out.get("r_s")), _ptr(out.get("Ir")),
                                   _ptr(out.get("Imax")), _ptr(out.get("Nmax")), _ptr(out.get("zero_mask")),
                                   _stream_handle())
    _lib.check(rc, "kgeo_equatorial_image")
    return out


def equatorial_image_vjp_device(a, r_o, th_o, alpha, beta, grad_image=None, mbar_max=MBAR_MAX, p=2.0,
                                want_jac=False):
    """(grad3 tensor [3], jac tensor [3, npix] or None): sum_i grad_i dI_i/d(a, th_o, p)."""
    lib = _lib.load()
    dev = _device()
    al = _as_device_f64(alpha, dev).reshape(-1)
    be = _as_device_f64(beta, dev).reshape(-1)
    n = al.numel()
    gi = None if grad_image is None else _as_device_f64(grad_image, dev).reshape(-1)
    jac = torch.empty((3, n), device=dev, dtype=torch.float64) if want_jac else None
    grad3 = torch.empty(3, device=dev, dtype=torch.float64)
    wsb = int(lib.kgeo_equatorial_vjp_workspace_bytes(n))
    ws = torch.empty(wsb, device=dev, dtype=torch.uint8)
    prm = _params(a, r_o, th_o, n)
    rc = lib.kgeo_equatorial_image_vjp(prm, _ptr(al), _ptr(be), int(mbar_max), float(p), _ptr(gi), _ptr(jac),
                                       _ptr(grad3), _ptr(ws), wsb, _stream_handle())
    _lib.check(rc, "kgeo_equatorial_image_vjp")
    return grad3, jac


class EquatorialImageFn(torch.autograd.Function):
    """image = EquatorialImageFn.apply(theta, alpha, beta, r_o, mbar_max) with
    theta = (a, th_o, p) a float64 tensor of shape (3,)."""

    @staticmethod
    def forward(ctx, theta, alpha, beta, r_o, mbar_max):
        a, th_o, p = (float(x) for x in theta.detach().cpu())
        out = equatorial_image_device(a, r_o, th_o, alpha, beta, mbar_max, p, want_parts=False)
        ctx.save_for_backward(theta, alpha, beta)
        ctx.r_o = r_o
        ctx.mbar_max = mbar_max
        return out["image"]

    @staticmethod
    def backward(ctx, grad_image):
        theta, alpha, beta = ctx.saved_tensors
        a, th_o, p = (float(x) for x in theta.detach().cpu())
        grad3, _ = equatorial_image_vjp_device(a, ctx.r_o, th_o, alpha, beta, grad_image.contiguous(),
                                               ctx.mbar_max, p)
        return grad3.to(theta.device, theta.dtype), None, None, None, None


def critical_curve(a, th0, n=100000):
    """Screen-plane critical curve (alpha, beta), as equatorial_lensing.py:264-293
    (parametrised by the photon-orbit radius r in [r_-, r_+], eq. 38-40 of GL19b)."""
    rminus = 2 * (1 + np.cos(2. * np.arccos(-a) / 3.))
    rplus = 2 * (1 + np.cos(2. * np.arccos(a) / 3.))
    r = np.linspace(rminus, rplus, n)
    with np.errstate(invalid="ignore", divide="ignore"):
        lam = -(r ** 3 + a ** 2 + r * (a ** 2) - 3 * r ** 2) / (a * r - a)
        eta = (r ** 2) * (3 * r ** 2 + a ** 2 - lam ** 2) / (r ** 2 - a ** 2)
        alpha = -lam / np.sin(th0)
        beta = np.sqrt(eta + (a * np.cos(th0)) ** 2 - (lam / np.tan(th0)) ** 2)
    bad = np.isnan(alpha) + np.isnan(beta)
    alpha, beta = alpha[~bad], beta[~bad]
    return np.hstack((alpha, np.flip(alpha))), np.hstack((beta, -np.flip(beta)))


def exclusion_mask(a, th_o, alpha, beta, r_o=1.0e4, p=2.0, mbar_max=MBAR_MAX, crit_dist=1.0e-3, fd_step=1.0e-6):
    """True for pixels kept in the loss: farther than crit_dist from the
    critical curve and with the same N_max and subring masks at every
    (a, th_o, p) -+ fd_step, so the image is smooth across the
    finite-difference stencil.  Returns (keep, n_near_curve, n_discontinuous)."""
    from scipy.spatial import cKDTree
    alpha = np.asarray(alpha, float)
    beta = np.asarray(beta, float)
    ca, cb = critical_curve(a, th_o)
    d, _ = cKDTree(np.c_[ca, cb]).query(np.c_[alpha, beta])
    near = d < crit_dist
    base = np.array([a, th_o, p])
    ref = None
    disc = np.zeros(len(alpha), bool)
    for k in range(3):
        for sgn in (-1.0, 1.0):
            q = base.copy()
            q[k] += sgn * fd_step
            o = equatorial_image_device(q[0], r_o, q[1], alpha, beta, mbar_max, q[2])
            torch.cuda.current_stream().synchronize()
            key = (o["Nmax"].cpu().numpy(), o["zero_mask"].cpu().numpy())
            if ref is None:
                o0 = equatorial_image_device(a, r_o, th_o, alpha, beta, mbar_max, p)
                torch.cuda.current_stream().synchronize()
                ref = (o0["Nmax"].cpu().numpy(), o0["zero_mask"].cpu().numpy())
            disc |= (key[0] != ref[0]) | (key[1] != ref[1])
    keep = ~near & ~disc
    return keep, int(near.sum()), int((disc & ~near).sum())


def l2_loss(theta, alpha, beta, target, keep, r_o=1.0e4, mbar_max=MBAR_MAX):
    """||image(theta) - target||_2 over the kept pixels (differentiable in theta)."""
    img = EquatorialImageFn.apply(theta, alpha, beta, r_o, mbar_max)
    diff = (img - target) * keep
    return torch.sqrt(torch.sum(diff * diff))
