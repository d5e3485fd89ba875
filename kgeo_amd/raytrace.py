"""Drop-in ``raytrace_ana`` / ``coords_at_tau`` on the MI355X kernels.

Mirrors kgeo.kerr_raytracing_ana.raytrace_ana (kerr_raytracing_ana.py:46-122)
and coords_at_tau (:126-193): same keyword arguments and defaults, same
validation messages, and a ``Geodesics`` result with the reference's
attribute names and (ngeo, npix) float64 arrays.

Device memory and streams come from PyTorch; the arithmetic runs in the HIP
kernels of libkgeo_hip.so behind the C-ABI in include/kgeo.h.  The
``*_device`` variants keep everything on the GPU and return torch tensors.
"""

import numpy as np
import torch

from . import _lib
from .geodesics import Geodesics

# reference module defaults (kerr_raytracing_ana.py:21-29)
SPIN = 0.94
INC = 20 * np.pi / 180.
ROUT = 100000.
NGEO = 250
NPIX = 1000
MAXTAUFRAC = (1. - 1.e-10)
MINSPIN = 1.e-4
EP = 1.e-12

alpha_default = np.linspace(-6, 6, NPIX)
beta_default = 0 * alpha_default


def _device():
    if not torch.cuda.is_available():
        raise _lib.KgeoError("kgeo_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                             "there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


def _stream_handle():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(x):
    import ctypes
    return ctypes.c_void_p(int(x)) if x else None


def _ptr(t):
    return None if t is None else ctypes_void(t.data_ptr())


def validate(a, r_o, th_o):
    """Checks of kerr_raytracing_ana.py:58-66 with the reference's messages."""
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")
    if not (isinstance(r_o, float) and (r_o >= 100)):
        raise Exception("r_o should be a float > 100")
    if not (isinstance(th_o, float) and (0 < th_o < np.pi) and th_o != 0.5 * np.pi):
        raise Exception("th_o should be a float in range (0,pi/2) or (pi/2,pi)")


def _screen(alpha, beta):
    if not isinstance(alpha, np.ndarray):
        alpha = np.array([alpha]).flatten()
    if not isinstance(beta, np.ndarray):
        beta = np.array([beta]).flatten()
    if len(alpha) != len(beta):
        raise Exception("alpha, beta are different lengths!")
    return alpha, beta


def _minspin_check(a, th_o, alpha, beta):
    """The reference raises in th_integrate when |a| < MINSPIN and some eta <= 0
    survives the eta<0 -> EP fix, i.e. eta == 0 exactly (kerr_raytracing_ana.py:300-302)."""
    if np.abs(a) < MINSPIN:
        eta = (alpha ** 2 - a ** 2) * np.cos(th_o) ** 2 + beta ** 2
        eta[eta < 0] = EP
        if np.any(eta <= 0.):
            raise Exception("below MINSPIN but there are eta<0 points in th_integrate!")


def _as_device_f64(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev)


class RaytraceResult:
    """Device-side outputs of one raytrace (torch tensors on the GPU)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def raytrace_ana_device(a, r_o, th_o, alpha, beta, ngeo, maxtaufrac=MAXTAUFRAC, do_phi_and_t=True,
                        want=("tau", "t", "r", "th", "ph", "sig"), out=None):
    """Raytrace on the GPU and leave the results there.

    alpha, beta: 1-D tensors or arrays (float64).  Returns a RaytraceResult with
    (ngeo, npix) tensors tau, t, r, th, ph, sig (those in ``want``; t and ph
    are always produced, the theta pass parks its integrals in them), per-pixel
    rclass (int8), thclass (int8), nmax_eq (int32), tau_tot, status (int32).
    ``out`` may hold preallocated tensors of the same names.  Asynchronous on
    torch's current stream.  do_phi_and_t is accepted and ignored, as in the
    reference's raytrace_ana (kerr_raytracing_ana.py:97 always samples t and phi).
    """
    lib = _lib.load()
    dev = _device()
    al = _as_device_f64(alpha, dev).reshape(-1)
    be = _as_device_f64(beta, dev).reshape(-1)
    npix = al.numel()
    if be.numel() != npix:
        raise Exception("alpha, beta are different lengths!")
    out = dict(out or {})
    f64 = dict(device=dev, dtype=torch.float64)
    for k in ("tau", "t", "r", "th", "ph", "sig"):
        if (k in want or k in ("t", "ph")) and k not in out:
            out[k] = torch.empty((ngeo, npix), **f64)
    for k, dt in (("rclass", torch.int8), ("thclass", torch.int8), ("nmax_eq", torch.int32),
                  ("tau_tot", torch.float64), ("status", torch.int32)):
        if k not in out:
            out[k] = torch.empty(npix, device=dev, dtype=dt)
    wsb = _lib.workspace_bytes(npix, ngeo)
    ws = out.get("workspace")
    if ws is None or ws.numel() < wsb:
        ws = torch.empty(wsb + 256, device=dev, dtype=torch.uint8)
    p = _lib.make_params(a, r_o, th_o, npix, ngeo, maxtaufrac)
    rc = lib.kgeo_raytrace_ana(p, _ptr(al), _ptr(be), _ptr(out.get("tau")), _ptr(out.get("t")), _ptr(out.get("r")),
                               _ptr(out.get("th")), _ptr(out.get("ph")), _ptr(out.get("sig")),
                               _ptr(out["rclass"]), _ptr(out["thclass"]), _ptr(out["nmax_eq"]),
                               _ptr(out["tau_tot"]), _ptr(out["status"]), _ptr(ws), wsb, _stream_handle())
    _lib.check(rc, "kgeo_raytrace_ana")
    out["workspace"] = ws
    out["alpha"], out["beta"] = al, be
    return RaytraceResult(**out)


def coords_at_tau_device(a, r_o, th_o, alpha, beta, tau, do_phi_and_t=True):
    """Coordinates at caller-given Mino times tau (ntau, npix) or (npix,) on the GPU.
    Returns (sig, [t, r, th, ph]) as (ntau, npix) tensors.

    With do_phi_and_t=False the reference's th_integrate / r_integrate leave
    G_phi, G_t, I_+-, I_1, I_2 at zero (kerr_raytracing_ana.py:212-214,
    394-398), so coords_at_tau returns sigma = 0, phi = 0 and t = 4 I_0 =
    -4 tau (:186-188, :626-631); those values are returned here, with r and
    theta from the kernels (which always evaluate the full integrals)."""
    lib = _lib.load()
    dev = _device()
    al = _as_device_f64(alpha, dev).reshape(-1)
    be = _as_device_f64(beta, dev).reshape(-1)
    npix = al.numel()
    ta = _as_device_f64(tau, dev)
    if ta.dim() == 1:
        ta = ta.reshape(1, -1)
    if ta.shape[-1] != npix:
        raise Exception("tau has incompatible shape in coords_at_tau!")
    ntau = ta.shape[0]
    f64 = dict(device=dev, dtype=torch.float64)
    o = {k: torch.zeros((ntau, npix), **f64) for k in ("t", "r", "th", "ph", "sig")}
    status = torch.empty(npix, device=dev, dtype=torch.int32)
    wsb = _lib.workspace_bytes(npix)
    ws = torch.empty(wsb + 256, device=dev, dtype=torch.uint8)
    p = _lib.make_params(a, r_o, th_o, npix, ntau, MAXTAUFRAC)
    rc = lib.kgeo_coords_at_tau(p, _ptr(al), _ptr(be), _ptr(ta), ntau, _ptr(o["t"]), _ptr(o["r"]), _ptr(o["th"]),
                                _ptr(o["ph"]), _ptr(o["sig"]), _ptr(status), _ptr(ws), wsb, _stream_handle())
    _lib.check(rc, "kgeo_coords_at_tau")
    if not do_phi_and_t:
        o["sig"].zero_()
        o["ph"].zero_()
        torch.mul(ta, -4.0, out=o["t"])
    return o["sig"], [o["t"], o["r"], o["th"], o["ph"]]


def raytrace_ana(a=SPIN, observer_coords=[0, ROUT, INC, 0], image_coords=[alpha_default, beta_default],
                 ngeo=NGEO, maxtaufrac=MAXTAUFRAC, do_phi_and_t=True, savedata=False, plotdata=True):
    """Analytically raytrace over the full range of Mino times (kerr_raytracing_ana.py:46-122).

    As in the reference, t, phi and sigma are always computed (the reference
    passes do_phi_and_t=True to coords_at_tau unconditionally, :97) and
    do_phi_and_t only gates saving and plotting.  Returns Geodesics with
    numpy arrays.
    """
    [_, r_o, th_o, _] = observer_coords
    [alpha, beta] = image_coords
    validate(a, r_o, th_o)
    alpha, beta = _screen(alpha, beta)
    _minspin_check(a, th_o, alpha, beta)
    res = raytrace_ana_device(a, r_o, th_o, alpha, beta, ngeo, maxtaufrac, do_phi_and_t=True)
    torch.cuda.current_stream().synchronize()
    host = {k: getattr(res, k).cpu().numpy() for k in ("tau", "t", "r", "th", "ph", "sig")}
    geos = Geodesics(a, observer_coords, [alpha, beta], host["tau"], host["sig"],
                     [host["t"], host["r"], host["th"], host["ph"]])
    geos._device_labels = {k: getattr(res, k).cpu().numpy() for k in ("rclass", "thclass", "nmax_eq", "tau_tot",
                                                                      "status")}
    if savedata and do_phi_and_t:
        try:
            geos.savegeos()
        except Exception:
            print("Error saving to file!")
    if plotdata and do_phi_and_t:
        try:
            geos.plotgeos()
        except Exception:
            print("Error plotting data!")
    return geos


def coords_at_tau(a, observer_coords, image_coords, tau, do_phi_and_t=True):
    """Coordinates x^mu at given Mino times (kerr_raytracing_ana.py:126-193)."""
    [_, r_o, th_o, _] = observer_coords
    [alpha, beta] = image_coords
    validate(a, r_o, th_o)
    alpha, beta = _screen(alpha, beta)
    tau = np.asarray(tau, dtype=np.float64)
    if len(tau.shape) == 1:
        tau = tau.reshape(1, tau.shape[0])
    if not (tau.shape[1] == len(alpha)):
        raise Exception("tau has incompatible shape in coords_at_tau!")
    _minspin_check(a, th_o, alpha, beta)
    sig, (t, r, th, ph) = coords_at_tau_device(a, r_o, th_o, alpha, beta, tau, do_phi_and_t=do_phi_and_t)
    torch.cuda.current_stream().synchronize()
    return sig.cpu().numpy(), [t.cpu().numpy(), r.cpu().numpy(), th.cpu().numpy(), ph.cpu().numpy()]
