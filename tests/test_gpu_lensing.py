"""GPU parity of the lensed-curve kernels (kgeo_rho_of_req,
kgeo_critical_curve) against the oracle (equatorial_lensing.py:198-423),
through the C-ABI: many curves in one launch, the drop-ins' return values
and errors, and a small curve library in the reference's file format."""

import os

import numpy as np
import pytest

from oracle import kgeo_oracle as O

from .test_hostcheck_lensing import roots_agree

pytestmark = pytest.mark.gpu

CURVES = [(0.5, 30.0, 6.0, 0), (0.9, 60.0, None, 1), (0.3, 17.0, 10.0, 2), (-0.7, 80.0, 4.0, 0),
          (0.99, 85.0, None, 0), (0.0, 45.0, 5.0, 1), (0.7, 5.0, 20.0, 0), (-0.3, 120.0, 6.0, 1)]


def _req(a, req):
    return (1 + np.sqrt(1 - a * a)) if req is None else req


def test_many_curves_in_one_launch_match_oracle():
    import torch
    from kgeo_amd.lensing import rho_of_req_device
    vp = np.linspace(-180, 179, 72) * np.pi / 180
    spins = [c[0] for c in CURVES]
    incs = [np.deg2rad(c[1]) for c in CURVES]
    reqs = [_req(c[0], c[2]) for c in CURVES]
    mbars = [c[3] for c in CURVES]
    rho, status = rho_of_req_device(spins, incs, reqs, mbars, vp)
    torch.cuda.synchronize()
    rho, status = rho.cpu().numpy(), status.cpu().numpy()
    assert np.all(status == 0)
    worst = 0.0
    for k, (a, thd, req, mbar) in enumerate(CURVES):
        th = np.deg2rad(thd)
        _, ref, _, _ = O.rho_of_req(a, th, reqs[k], mbar, vp)
        assert not roots_agree(a, th, reqs[k], mbar, vp, rho[k], ref), (a, thd, mbar)
        worst = max(worst, float(np.max(np.abs(rho[k] - ref))))
    print(f"\n[rho_of_req] {len(CURVES)} curves x {len(vp)} angles: largest |rho - oracle| {worst:.1e}")


def test_dropins_and_errors():
    from kgeo_amd.lensing import critical_curve, rho_of_req, rho_of_req_single
    a, th = 0.6, np.deg2rad(40.0)
    vp = np.linspace(-180, 179, 30) * np.pi / 180
    v, rho, al, be = rho_of_req(a, th, 7.0, 0, vp)
    _, ref, ral, rbe = O.rho_of_req(a, th, 7.0, 0, vp)
    assert np.allclose(rho, ref, rtol=0, atol=1e-10) and np.allclose(al, ral, atol=1e-10)
    assert abs(rho_of_req_single(a, th, 7.0, float(vp[3])) - ref[3]) < 1e-10
    with pytest.raises(ValueError, match="different signs"):
        rho_of_req(a, th, 0.5, 0, vp)
    for aa, thd in ((0.5, 30.0), (0.94, 17.0), (-0.7, 70.0)):
        ga, gb = critical_curve(aa, np.deg2rad(thd), n=20000)
        oa, ob = O.critical_curve(aa, np.deg2rad(thd), n=20000)
        assert ga.shape == oa.shape
        assert np.max(np.abs(ga - oa) / np.maximum(np.abs(oa), 1)) < 1e-12
        assert np.max(np.abs(gb - ob) / np.maximum(np.abs(ob), 1)) < 1e-10


def test_library_files(tmp_path):
    from kgeo_amd.lensing import generate_library
    vp = np.linspace(-180, 179, 36) * np.pi / 180
    out = str(tmp_path / "lib")
    files = generate_library("rh", outpath=out, spins=[0.3, 0.8], incs=[20.0, 60.0], mmax=2, varphis=vp)
    assert len(files) == 4
    f = os.path.join(out, "a80_i60_rh.txt")
    assert f in files
    tab = np.loadtxt(f)
    assert tab.shape == (36, 4) and np.allclose(tab[:, 0], vp)
    a, th = 0.8, np.deg2rad(60.0)
    for m in range(3):
        _, ref, _, _ = O.rho_of_req(a, th, 1 + np.sqrt(1 - a * a), m, vp)
        assert not roots_agree(a, th, 1 + np.sqrt(1 - a * a), m, vp, tab[:, 1 + m], ref, tol=1e-10)
    # existing files are kept (the reference's resume)
    assert generate_library("rh", outpath=out, spins=[0.3, 0.8], incs=[20.0, 60.0], mmax=2, varphis=vp) == []
