"""ctypes binding of the kgeo C-ABI (include/kgeo.h).

The HIP library ``libkgeo_hip.so`` is built in-tree next to this file by
``__graft_entry__.build()`` (or ``python -m kgeo_amd.build``).  There is no
CPU fallback: if the library is missing, every entry point raises.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# must equal KGEO_ABI_VERSION of include/kgeo.h
ABI_VERSION = 4
LIB_PATH = os.path.join(_HERE, "libkgeo_hip.so")

KGEO_DO_PHI_T = 1

VJP_NOALLOC_MAX_NPIX = 8388608   # KGEO_VJP_NOALLOC_MAX_NPIX

ST_NONFINITE = 1
ST_DEGENERATE_ROOTS = 2
ST_MINSPIN_VORTICAL = 4
ST_LADDER_CAP = 8
ST_LABEL_TIE = 16

c_double_p = ctypes.POINTER(ctypes.c_double)


class RayParams(ctypes.Structure):
    """kgeo_ray_params (include/kgeo.h)."""

    _fields_ = [
        ("a", ctypes.c_double), ("r_o", ctypes.c_double), ("th_o", ctypes.c_double),
        ("maxtaufrac", ctypes.c_double),
        ("a2", ctypes.c_double), ("sin_th", ctypes.c_double), ("cos_th", ctypes.c_double),
        ("cos2_th", ctypes.c_double), ("rplus", ctypes.c_double), ("rminus", ctypes.c_double),
        ("npix", ctypes.c_int64), ("ngeo", ctypes.c_int64),
        ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
    ]


class BatchTuning(ctypes.Structure):
    """kgeo_batch_tuning (include/kgeo.h)."""

    _fields_ = [("samples_per_item", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 2)]


class BfieldModel(ctypes.Structure):
    """kgeo_bfield_model (include/kgeo.h)."""

    _fields_ = [("kind", ctypes.c_int32), ("secondorder_only", ctypes.c_int32), ("usemono", ctypes.c_int32),
                ("is_above", ctypes.c_int32), ("C", ctypes.c_double), ("Cr", ctypes.c_double),
                ("Cvert", ctypes.c_double), ("Cph", ctypes.c_double), ("omegafac", ctypes.c_double),
                ("p", ctypes.c_double), ("n_I", ctypes.c_double), ("p_val", ctypes.c_double),
                ("table", ctypes.c_void_p), ("table_n", ctypes.c_int64)]


class VelocityModel(ctypes.Structure):
    """kgeo_velocity_model (include/kgeo.h)."""

    _fields_ = [("kind", ctypes.c_int32), ("retrograde", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in ("fac_subkep", "beta_phi", "beta_r", "gelles_beta", "gelles_chi", "ell_isco",
                                       "vr_isco", "p1", "p2", "dd", "nu_parallel")] + [
        ("drift_field", BfieldModel), ("table", ctypes.c_void_p), ("table_n", ctypes.c_int64)]


class EmissivityModel(ctypes.Structure):
    """kgeo_emissivity_model (include/kgeo.h)."""

    _fields_ = [("kind", ctypes.c_int32), ("bfield_model", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in ("emiscut_in", "emiscut_out", "p1", "p2", "mu_ring", "gamma_off", "sigma", "Rb",
                                       "ne0", "Te0", "B0", "alpha_n", "alpha_T", "alpha_B")]


class EmissionModel(ctypes.Structure):
    """kgeo_emission_model (include/kgeo.h)."""

    _fields_ = [("velocity", VelocityModel), ("bfield", BfieldModel), ("emissivity", EmissivityModel),
                ("polarization", ctypes.c_int32), ("pathlength", ctypes.c_int32), ("specind", ctypes.c_double),
                ("diskangle", ctypes.c_double), ("nu_obs", ctypes.c_double)]


class KgeoError(RuntimeError):
    pass


_lib = None

# name -> (restype, argtypes)
_PROTOS = {
    "kgeo_abi_version": (ctypes.c_int, []),
    "kgeo_last_error": (ctypes.c_char_p, []),
    "kgeo_fill_params": (None, [ctypes.POINTER(RayParams)]),
    "kgeo_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int64]),
    "kgeo_sample_chunks": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64]),
    "kgeo_raytrace_ana": (ctypes.c_int, [ctypes.POINTER(RayParams)] + [ctypes.c_void_p] * 14
                          + [ctypes.c_size_t, ctypes.c_void_p]),
    "kgeo_coords_at_tau": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 7
                           + [ctypes.c_size_t, ctypes.c_void_p]),
    "kgeo_plan": (ctypes.c_int, [ctypes.POINTER(RayParams)] + [ctypes.c_void_p] * 8
                  + [ctypes.c_size_t, ctypes.c_void_p]),
    "kgeo_sample": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_void_p, ctypes.c_int64]
                    + [ctypes.c_void_p] * 8 + [ctypes.c_size_t, ctypes.c_void_p]),
    "kgeo_ellip_pi": (ctypes.c_int, [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_void_p]),
    "kgeo_equatorial_image": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_double] + [ctypes.c_void_p] * 9),
    "kgeo_equatorial_vjp_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int64]),
    "kgeo_equatorial_image_vjp": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_int, ctypes.c_double] + [ctypes.c_void_p] * 4
                                  + [ctypes.c_size_t, ctypes.c_void_p]),
    "kgeo_equatorial_last_error": (ctypes.c_char_p, []),
    "kgeo_rho_of_req": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64]
                        + [ctypes.c_void_p] * 3),
    "kgeo_critical_curve": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_void_p, ctypes.c_int64]
                            + [ctypes.c_void_p] * 3),
    "kgeo_equatorial_image_vjp_noalloc": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_void_p,
                                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
                                          + [ctypes.c_void_p] * 4),
    "kgeo_equatorial_stokes": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int32, ctypes.POINTER(EmissionModel)]
                               + [ctypes.c_void_p] * 12),
    "kgeo_model_eval": (ctypes.c_int, [ctypes.POINTER(EmissionModel), ctypes.c_double, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 10),
    "kgeo_emission_last_error": (ctypes.c_char_p, []),
    "kgeo_crossings": (ctypes.c_int, [ctypes.POINTER(RayParams)] + [ctypes.c_void_p] * 3
                       + [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "kgeo_turnings": (ctypes.c_int, [ctypes.POINTER(RayParams)] + [ctypes.c_void_p] * 3
                      + [ctypes.c_int64] + [ctypes.c_void_p] * 4),
    "kgeo_batch_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(RayParams), ctypes.c_int32]),
    "kgeo_batch_layout": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "kgeo_batch_plan": (ctypes.c_int, [ctypes.POINTER(RayParams), ctypes.c_int32, ctypes.POINTER(BatchTuning)]
                        + [ctypes.c_void_p] * 12 + [ctypes.c_size_t, ctypes.c_void_p,
                                                    ctypes.POINTER(ctypes.c_void_p)]),
    "kgeo_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "kgeo_batch_items": (ctypes.c_int64, [ctypes.c_void_p]),
    "kgeo_batch_progress": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "kgeo_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "kgeo_batch_last_error": (ctypes.c_char_p, []),
}


def exported_symbols():
    """Names the library must export (every function declared in include/kgeo.h)."""
    return sorted(_PROTOS)


def load():
    """Load libkgeo_hip.so (raises KgeoError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KgeoError(f"HIP library {LIB_PATH} is missing: run __graft_entry__.build() "
                        "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.kgeo_abi_version() != ABI_VERSION:
        raise KgeoError("libkgeo_hip.so ABI version mismatch")
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        lib = load()
        if what in ("kgeo_equatorial_stokes", "kgeo_model_eval"):
            err = lib.kgeo_emission_last_error
        elif what.startswith("kgeo_equatorial") or what in ("kgeo_rho_of_req", "kgeo_critical_curve"):
            err = lib.kgeo_equatorial_last_error
        elif what.startswith("kgeo_batch"):
            err = lib.kgeo_batch_last_error
        else:
            err = lib.kgeo_last_error
        msg = err().decode(errors="replace")
        raise KgeoError(f"{what} failed ({rc}): {msg}")


def make_params(a, r_o, th_o, npix, ngeo, maxtaufrac, do_phi_t=True):
    """kgeo_ray_params with the derived scalars computed by numpy exactly as the
    reference writes them (kerr_raytracing_ana.py:76-77, :390-391), so that the
    device arithmetic on (alpha, beta) reproduces the reference bit for bit."""
    p = RayParams()
    p.a = float(a)
    p.r_o = float(r_o)
    p.th_o = float(th_o)
    p.maxtaufrac = float(maxtaufrac)
    p.a2 = float(a ** 2)
    p.sin_th = float(np.sin(th_o))
    p.cos_th = float(np.cos(th_o))
    p.cos2_th = float(np.cos(th_o) ** 2)
    p.rplus = float(1 + np.sqrt(1 - a ** 2))
    p.rminus = float(1 - np.sqrt(1 - a ** 2))
    p.npix = int(npix)
    p.ngeo = int(ngeo)
    p.flags = KGEO_DO_PHI_T if do_phi_t else 0
    p.reserved = 0
    return p


def workspace_bytes(npix, ngeo=0):
    """Device workspace bytes for npix pixels on a linspace grid of ngeo Mino
    times (0: the plan alone, as caller-given Mino times need)."""
    return int(load().kgeo_workspace_bytes(int(npix), int(ngeo)))
