"""Runs the CPU instantiation of the device source under AddressSanitizer and
UndefinedBehaviorSanitizer (kgeo_amd/libkgeo_hostcheck_san.so) over launch
shapes and regimes that exercise every pass, queue and record index, and
compares its outputs with the plain CPU build.  Started by
tests/test_sanitized_host.py in a child process whose LD_PRELOAD holds the
sanitizer runtime; any sanitizer report aborts it with a non-zero status."""

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from tests.kgeo_testlib import Host  # noqa: E402


def run():
    plain = Host()
    san = Host.__new__(Host)
    san.lib = ctypes.CDLL(os.path.join(ROOT, "kgeo_amd", "libkgeo_hostcheck_san.so"))
    rng = np.random.default_rng(11)
    cases = [(0.94, 17.0, 8.0), (-0.9, 80.0, 7.0), (0.5, 140.0, 9.0), (0.0, 60.0, 6.0), (0.999, 89.0, 5.5)]
    checked = 0
    for a, th_deg, w in cases:
        th = th_deg * np.pi / 180
        npix = int(rng.integers(37, 90))
        al, be = rng.uniform(-w, w, npix), rng.uniform(-w, w, npix)
        al[:3], be[:3] = [0.0, 1e-9, -w], [w, 0.0, 1e-9]          # axis and near-axis pixels
        for ngeo in (3, 4, 17, 33, 100, 257):
            for chunks in (1, 3, 7):
                for do_phi_t in (1, 0):
                    outs = []
                    for h in (plain, san):
                        o = h.raytrace_loops(a, 1e4, th, al, be, ngeo, chunks=chunks, do_phi_t=do_phi_t)
                        outs.append(o)
                    for k in ("r", "th", "ph", "t"):
                        if not do_phi_t and k in ("ph", "t"):
                            continue
                        # the two builds inline and contract differently (the sanitizers change the
                        # optimiser's choices); they must agree to well within the parity bar, except
                        # at the last sample of horizon-bound rays, where r - r_+ is ill-conditioned
                        # and a log of it may be NaN in one build and finite in the other
                        x, y = outs[0][k][:-1], outs[1][k][:-1]
                        assert np.array_equal(np.isnan(x), np.isnan(y)), (a, th_deg, ngeo, chunks, k)
                        ok = np.isfinite(x)
                        assert np.array_equal(x[~ok & ~np.isnan(x)], y[~ok & ~np.isnan(x)]), (a, ngeo, k)
                        scale = np.maximum(np.abs(x), 1.0) if k in ("r", "th") else \
                            np.maximum(np.nanmax(np.where(ok, np.abs(x), 0), axis=0, keepdims=True), 1.0)
                        err = np.abs(np.where(ok, x - y, 0)) / scale
                        assert err.max() < 1e-6, (a, th_deg, ngeo, chunks, do_phi_t, k, float(err.max()))
                    checked += 1
        # the other device headers: counts, equatorial image, lensed curves, emission
        tau = np.abs(rng.normal(size=(5, npix))) * 3
        for h in (plain, san):
            h.counts(a, 1e4, th, al, be, tau)
    print(f"sanitized runs clean: {checked} pass configurations", flush=True)


if __name__ == "__main__":
    run()
