"""Writes the committed fixtures under tests/golden/.

  ellip_pi_mpmath.npz   Pi(n; phi | m) truth from mpmath at 30 digits
                        (real part = Cauchy principal value for n > 1) on a
                        grid spanning the reference's regimes: m < 0 (theta
                        integrals), 0 < m < 1 (radial), n < 0, 0 < n < 1,
                        n > 1, |phi| up to several periods.  Also F and E.
  oracle_c1_strat.npz   oracle outputs (kgeo_oracle.raytrace_ana) for a
                        class-stratified subset of the 64x64, ngeo=100
                        workload, to detect drift of the oracle itself.
  ipole_midplane.json   (hand-copied data) the 15 first-midplane-crossing
                        known answers of the reference's own test,
                        tests/test_geodesics_kerr_ana.py:54-70, rtol 1e-3.

Only mpmath, numpy and the oracle are used; the reference is not imported.
Run from the repository root:  python tests/golden/make_golden.py
"""

import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import kgeo_oracle as O  # noqa: E402


def ellip_fixture():
    mp.mp.dps = 30
    rng = np.random.default_rng(12345)
    ns, phis, ms = [], [], []
    for nlo, nhi in ((-3.0, 0.0), (0.0, 1.0), (1.0, 4.0)):
        for mlo, mhi in ((-50.0, 0.0), (0.0, 0.999)):
            k = 40
            ns.append(rng.uniform(nlo, nhi, k))
            ms.append(rng.uniform(mlo, mhi, k))
            phis.append(rng.uniform(-12.0, 12.0, k))
    # the reference's own print-check arguments (scipy_ellip_binding.py:241-246)
    ref_phi = [.026, .1, np.pi / 2., 2.37, 4.1, 6., 7.7, 14.3, 215.]
    ref_phi = ref_phi + [-p for p in ref_phi]
    ns.append(np.full(len(ref_phi), .8))
    ms.append(np.full(len(ref_phi), -.33))
    phis.append(np.array(ref_phi))
    n = np.concatenate(ns)
    m = np.concatenate(ms)
    phi = np.concatenate(phis)
    P = np.array([float(mp.re(mp.ellippi(float(a), float(b), float(c)))) for a, b, c in zip(n, phi, m)])
    F = np.array([float(mp.ellipf(float(b), float(c))) for b, c in zip(phi, m)])
    E = np.array([float(mp.ellipe(float(b), float(c))) for b, c in zip(phi, m)])
    np.savez_compressed(os.path.join(HERE, "ellip_pi_mpmath.npz"), n=n, phi=phi, m=m, Pi=P, F=F, E=E)


def oracle_fixture():
    a, th_o, r_o, ngeo = 0.94, 17.0 * np.pi / 180.0, 1.0e4, 100
    al = np.linspace(-8, 8, 64)
    A, B = np.meshgrid(al, al)
    A, B = A.flatten(), B.flatten()
    lam, eta = O.conserved_fixed(a, th_o, A, B)
    _, _, _, _, rclass = O.radial_roots(a, lam, eta)
    rng = np.random.default_rng(7)
    pick = []
    for c in (1, 3, 4):
        idx = np.where(rclass == c)[0]
        pick.append(rng.choice(idx, min(len(idx), 24), replace=False))
    vort = np.where(eta < 0)[0]
    pick.append(rng.choice(vort, min(len(vort), 12), replace=False))
    idx = np.unique(np.concatenate(pick))
    out = O.raytrace_ana(a, r_o, th_o, A[idx], B[idx], ngeo)
    np.savez_compressed(os.path.join(HERE, "oracle_c1_strat.npz"), idx=idx, alpha=A[idx], beta=B[idx],
                        a=a, th_o=th_o, r_o=r_o, ngeo=ngeo,
                        **{k: out[k] for k in ("tau", "t", "r", "th", "ph", "sig", "rclass", "thclass", "tau_tot")})


if __name__ == "__main__":
    ellip_fixture()
    oracle_fixture()
    print("fixtures written")
