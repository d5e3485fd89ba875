"""The C-ABI library: loads without a GPU, exports every function declared
in include/kgeo.h, and rejects bad arguments with error codes (host-side
validation only; nothing here launches a kernel)."""

import ctypes
import os
import re
import subprocess

import pytest

from kgeo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "kgeo.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kgeo_[a-z_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert _declared() == _lib.exported_symbols()


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "run python -m kgeo_amd.build"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing


def test_load_and_version():
    lib = _lib.load()
    assert lib.kgeo_abi_version() == 4
    assert _lib.workspace_bytes(1000) == 256 + 1000 * _plan_slots() * 8


def test_workspace_holds_the_walk_tables():
    """On a linspace grid of ngeo > 2 samples the workspace adds the record
    tables of the anchor pass (theta: ceil(ngeo/32) + 3 records of 6 fields;
    radial: ceil(ngeo/16) + 3 records of 13 fields), the fixup queue (one
    8-byte item per 8 samples) and the repair queue (one item per 256-sample
    block), per pixel."""
    n = 1000
    plan = _lib.workspace_bytes(n)
    assert _lib.workspace_bytes(n, 2) == plan
    assert _lib.workspace_bytes(n, 500) == plan + n * 8 * (19 * 6 + 35 * 13 + 63 + 2)
    assert _lib.workspace_bytes(n, 33) == plan + n * 8 * (5 * 6 + 6 * 13 + 5 + 1)


def _plan_slots():
    lib = _lib.load()
    return (lib.kgeo_workspace_bytes(1, 0) - 256) // 8


def test_fill_params_matches_numpy_expressions():
    import numpy as np
    lib = _lib.load()
    p = _lib.RayParams()
    p.a, p.th_o = 0.94, 17 * np.pi / 180
    lib.kgeo_fill_params(ctypes.byref(p))
    q = _lib.make_params(0.94, 1e4, 17 * np.pi / 180, 1, 1, 1.0)
    for f in ("a2", "sin_th", "cos_th", "cos2_th", "rplus", "rminus"):
        assert abs(getattr(p, f) - getattr(q, f)) <= 2e-16 * max(1, abs(getattr(q, f))), f


@pytest.mark.parametrize("bad", ["null_params", "neg_npix", "no_workspace", "small_workspace", "ngeo0", "no_phi_t"])
def test_argument_errors(bad):
    lib = _lib.load()
    p = _lib.make_params(0.5, 1e4, 0.5, 10, 10, 1.0)
    buf = (ctypes.c_char * 8192)()
    addr = (ctypes.addressof(buf) + 255) // 256 * 256
    ws, wsb = ctypes.c_void_p(addr), 4096
    fake = ctypes.c_void_p(addr)   # never dereferenced: validation fails first
    pp = ctypes.byref(p)
    if bad == "null_params":
        pp = None
    elif bad == "neg_npix":
        p.npix = -1
    elif bad == "no_workspace":
        ws = None
    elif bad == "small_workspace":
        wsb = 16
    elif bad == "ngeo0":
        p.ngeo = 0
    elif bad == "no_phi_t":
        p.flags = 0
    rc = lib.kgeo_raytrace_ana(pp, fake, fake, None, None, None, None, None, None, None, None, None, None, None,
                               ws, wsb, None)
    assert rc < 0
    assert lib.kgeo_last_error()


def test_sizes_beyond_the_queue_encoding_are_refused():
    """One call handles npix < 2^31, ngeo < 2^31 and npix * ceil(ngeo / 8)
    fixup items within 32 bits (the queue's item and counter widths); larger
    jobs are refused before any device work, with a message to split them."""
    lib = _lib.load()
    buf = (ctypes.c_char * 8192)()
    fake = ctypes.c_void_p((ctypes.addressof(buf) + 255) // 256 * 256)
    p = _lib.make_params(0.5, 1e4, 0.5, 1 << 31, 10, 1.0)
    rc = lib.kgeo_raytrace_ana(ctypes.byref(p), fake, fake, None, fake, None, None, fake, None, None, None, None,
                               None, None, fake, 1 << 40, None)
    assert rc == -1 and b"split the pixels" in lib.kgeo_last_error()
    p = _lib.make_params(0.5, 1e4, 0.5, 1 << 24, 1 << 12, 1.0)   # 2^24 * 2^9 = 2^33 items
    rc = lib.kgeo_sample(ctypes.byref(p), None, 1 << 12, None, fake, fake, fake, fake, None, None, fake, 1 << 50,
                         None)
    assert rc == -1 and b"2^32" in lib.kgeo_last_error()


def test_every_raytrace_entry_requires_phi_t():
    """t, phi and sigma are always sampled (the reference's raytrace_ana does,
    kerr_raytracing_ana.py:97): requests without KGEO_DO_PHI_T are refused
    before any device work, by every entry point that samples rays."""
    lib = _lib.load()
    p = _lib.make_params(0.5, 1e4, 0.5, 10, 10, 1.0, do_phi_t=False)
    buf = (ctypes.c_char * 8192)()
    fake = ctypes.c_void_p((ctypes.addressof(buf) + 255) // 256 * 256)
    rcs = [
        lib.kgeo_plan(ctypes.byref(p), fake, fake, None, None, None, None, None, fake, 1 << 40, None),
        lib.kgeo_sample(ctypes.byref(p), None, 10, None, fake, fake, fake, fake, None, None, fake, 1 << 40, None),
        lib.kgeo_coords_at_tau(ctypes.byref(p), fake, fake, fake, 10, fake, fake, fake, fake, None, None, fake,
                               1 << 40, None),
    ]
    assert all(rc == -1 for rc in rcs), rcs
    assert b"KGEO_DO_PHI_T" in lib.kgeo_last_error()
    imgs = (_lib.RayParams * 2)(_lib.make_params(0.5, 1e4, 0.5, 64, 10, 1.0),
                                _lib.make_params(0.5, 1e4, 0.5, 64, 10, 1.0, do_phi_t=False))
    assert lib.kgeo_batch_workspace_bytes(imgs, 2) == 0
    h = ctypes.c_void_p()
    rc = lib.kgeo_batch_plan(imgs, 2, None, fake, fake, fake, fake, fake, fake, None, None, None, None, None,
                             fake, 1 << 40, fake, ctypes.byref(h))
    assert rc == -1 and b"KGEO_DO_PHI_T" in lib.kgeo_batch_last_error()


def test_turnings_argument_checks():
    """n_angular_turnings accepts th_o in [0, pi/2] only (kerr_raytracing_utils.py:690-691);
    the radial count alone has no such limit (checked before any launch)."""
    lib = _lib.load()
    buf = (ctypes.c_char * 64)()
    fake = ctypes.c_void_p(ctypes.addressof(buf))
    p = _lib.make_params(0.5, 1e4, 2.0, 4, 1, 1.0)
    assert lib.kgeo_turnings(ctypes.byref(p), fake, fake, fake, 1, fake, None, None, None) == -1
    assert b"(0,pi/2]" in lib.kgeo_last_error()
    assert lib.kgeo_turnings(ctypes.byref(p), fake, fake, fake, 1, None, fake, None, None) == -1
    # no pixels: accepted without touching the device
    p0 = _lib.make_params(0.5, 1e4, 2.0, 0, 1, 1.0)
    assert lib.kgeo_turnings(ctypes.byref(p0), None, None, None, 1, None, None, fake, None) == 0
    assert lib.kgeo_turnings(ctypes.byref(p), None, None, None, -1, None, None, None, None) == -1


def test_product_path_requires_gpu():
    # no CPU fallback: without a GPU the drop-in raises instead of computing
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from kgeo_amd.raytrace import raytrace_ana
    with pytest.raises(_lib.KgeoError):
        raytrace_ana(0.5, [0, 1e4, 0.5, 0], [[1.0], [2.0]], ngeo=4, plotdata=False)
