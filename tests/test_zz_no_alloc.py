"""The allocation-free gradient entry point (kgeo_equatorial_image_vjp_noalloc)
against the workspace one (kgeo_equatorial_image_vjp), through the C-ABI:
the same kernels and reduction order, so grad3 and the per-pixel Jacobian
must be bit-identical, on two calls in a row (the second smaller, so the
static scratch still holds the first call's partials) and on the C5 image
(256^2, a = 0.7, th_o = 40 deg, p = 2)."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R_O = 1.0e4


def _grid(w, n):
    al = np.linspace(-w, w, n)
    A, B = np.meshgrid(al, al)
    return A.flatten(), B.flatten()


def _both(a, th, p_emis, A, B, seed):
    import torch
    from kgeo_amd import _lib
    from kgeo_amd.raytrace import _ptr, _stream_handle
    lib = _lib.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    n = len(A)
    al = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
    be = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    w = torch.from_numpy(np.random.default_rng(seed).standard_normal(n)).to(dev)
    prm = _lib.make_params(a, R_O, th, n, 1, 1.0)
    out = {}
    for name in ("workspace", "noalloc"):
        jac = torch.full((3, n), float("nan"), device=dev, dtype=torch.float64)
        g3 = torch.full((3,), float("nan"), device=dev, dtype=torch.float64)
        if name == "workspace":
            wsb = int(lib.kgeo_equatorial_vjp_workspace_bytes(n))
            ws = torch.empty(wsb, device=dev, dtype=torch.uint8)
            rc = lib.kgeo_equatorial_image_vjp(prm, _ptr(al), _ptr(be), 2, p_emis, _ptr(w), _ptr(jac), _ptr(g3),
                                               _ptr(ws), wsb, _stream_handle())
        else:
            rc = lib.kgeo_equatorial_image_vjp_noalloc(prm, _ptr(al), _ptr(be), 2, p_emis, _ptr(w), _ptr(jac),
                                                       _ptr(g3), _stream_handle())
        _lib.check(rc, "kgeo_equatorial_image_vjp" + ("" if name == "workspace" else "_noalloc"))
        torch.cuda.synchronize()
        out[name] = (g3.cpu().numpy(), jac.cpu().numpy())
    return out


def test_noalloc_vjp_equals_workspace_vjp_on_two_calls():
    print("\n[no-alloc] call 1: 256^2 C5 image", flush=True)
    r1 = _both(0.7, np.deg2rad(40.0), 2.0, *_grid(12.0, 256), seed=1)
    print("[no-alloc] call 2: 97^2 (smaller, ragged) image", flush=True)
    r2 = _both(-0.4, np.deg2rad(70.0), 2.5, *_grid(9.0, 97), seed=2)
    for r in (r1, r2):
        (g_ws, j_ws), (g_na, j_na) = r["workspace"], r["noalloc"]
        assert np.all(np.isfinite(g_ws))
        assert np.array_equal(g_na, g_ws), (g_na, g_ws)
        assert np.array_equal(j_na, j_ws, equal_nan=True)
    assert not np.array_equal(r1["noalloc"][0], r2["noalloc"][0])


def test_noalloc_vjp_rejects_oversized_input_and_empty_image_gives_zero():
    import torch
    from kgeo_amd import _lib
    from kgeo_amd.raytrace import _ptr, _stream_handle
    lib = _lib.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    g3 = torch.full((3,), float("nan"), device=dev, dtype=torch.float64)
    big = _lib.make_params(0.5, R_O, 0.7, _lib.VJP_NOALLOC_MAX_NPIX + 1, 1, 1.0)
    rc = lib.kgeo_equatorial_image_vjp_noalloc(big, ctypes.c_void_p(16), ctypes.c_void_p(16), 2, 2.0, None, None,
                                               _ptr(g3), _stream_handle())
    assert rc == -1 and b"allocation-free" in lib.kgeo_equatorial_last_error()
    empty = _lib.make_params(0.5, R_O, 0.7, 0, 1, 1.0)
    _lib.check(lib.kgeo_equatorial_image_vjp_noalloc(empty, None, None, 2, 2.0, None, None, _ptr(g3),
                                                     _stream_handle()), "kgeo_equatorial_image_vjp_noalloc")
    torch.cuda.synchronize()
    assert np.array_equal(g3.cpu().numpy(), np.zeros(3))
