"""The sample loops' walk along the Mino-time grid (kgeo_amd/csrc/kg_step.hpp).

1. One step of the addition theorems (Jacobi sn/cn/dn, E, and the third-kind
   integral Pi, including the principal value for n > 1) against values
   computed independently: mpmath's Jacobi functions, scipy's E(phi|m) and the
   oracle's Pi (itself pinned to mpmath in test_oracle_elliptic.py).  A step
   that would cross a pole of the Pi integrand must be refused.
2. Whole rays: the walking loops (CPU build of the kernels' source) against
   the same loops forced to evaluate every sample directly (caller-given
   Mino times), over several spins, inclinations, ngeo and chunkings.
"""

import ctypes

import mpmath as mp
import numpy as np
import pytest
import scipy.special as sp

from oracle import kgeo_oracle as O

from .kgeo_testlib import HOSTCHECK, Host, errors, grid


def _lib():
    return ctypes.CDLL(HOSTCHECK)


def _truth(u, m, n):
    """(sn, cn, dn, (E - F)/m, Pi) at argument u, |u| < K(m), for parameter m."""
    mp.mp.dps = 30
    s, c, d = (float(mp.re(mp.ellipfun(f, u, m=m))) for f in ("sn", "cn", "dn"))
    phi = float(np.arcsin(s))  # |u| < K: am(u) = asin(sn u)
    E = float(sp.ellipeinc(phi, m))
    P = float(O.ellip_pi(np.array([n]), np.array([phi]), np.array([m]))[0])
    return np.array([s, c, d, (E - u) / m, P])


def _step(lib, at_u, at_v, u, v, m, n):
    out = np.zeros(5)
    P = lambda x: np.ascontiguousarray(x, dtype=float).ctypes.data_as(ctypes.c_void_p)
    lib.kgh_walk_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_double] * 4 + [ctypes.c_void_p]
    ok = lib.kgh_walk_step(P(at_u), P(at_v), u, v, m, n, P(out))
    return ok, out


@pytest.mark.parametrize("seed", range(4))
def test_one_step_against_independent_values(seed):
    lib = _lib()
    rng = np.random.default_rng(seed)
    taken = 0
    for _ in range(40):
        m = float(rng.choice([rng.uniform(-2.5, -0.05), rng.uniform(0.05, 0.97)]))
        n = float(rng.choice([rng.uniform(-3.0, 0.9), rng.uniform(1.2, 3.0)]))
        K = float(sp.ellipk(m)) if m > 0 else float(mp.ellipk(m))
        u = float(rng.uniform(-0.8, 0.8) * K)
        v = float(rng.uniform(-0.02, 0.02))
        want = _truth(u + v, m, n)
        ok, got = _step(lib, _truth(u, m, n), _truth(v, m, n), u, v, m, n)
        if not ok:
            # refused steps must be the ill-conditioned ones: a pole of the
            # integrand at or next to the step (n > 1, walked untransformed
            # only when m < 0)
            assert n > 1.0 and m < 0.0
            continue
        taken += 1
        scale = np.maximum(np.abs(want), 1.0)
        # sn, cn, dn, (E-F)/m to 2e-14; Pi to 5e-14 (for n > 1 it passes
        # through the m/n transformation and its log term on the way in and out)
        tol = np.array([2e-14, 2e-14, 2e-14, 2e-14, 5e-14])
        assert np.all(np.abs(got - want) / scale < tol), (m, n, u, v, got, want)
    assert taken >= 25


def test_step_across_a_pole():
    """n > 1: with 0 <= m < 1 the walk carries Pi(m/n) and the log term, so a
    step across the pole is taken and lands on the principal value; with
    m < 0 (walked untransformed) such a step is refused."""
    lib = _lib()
    for m, n, transformed in ((0.3, 1.6, True), (-0.4, 1.6, False)):
        # sn^2(u_p) = 1/n: the principal-value pole
        phi_p = np.arcsin(np.sqrt(1.0 / n))
        u_p = float(sp.ellipkinc(phi_p, m)) if m >= 0 else float(mp.ellipf(phi_p, m))
        for (u, v) in ((u_p - 0.004, 0.008), (u_p - 0.001, 0.0015), (u_p + 0.003, -0.006)):
            ok, got = _step(lib, _truth(u, m, n), _truth(v, m, n), u, v, m, n)
            if transformed:
                want = _truth(u + v, m, n)
                assert ok and np.all(np.abs(got - want) / np.maximum(np.abs(want), 1.0) < 1e-13), (u, v, got, want)
            else:
                assert not ok, (u, v)


CASES = [
    (0.94, 17.0, 8.0, 48, 500, 1),
    (0.94, 17.0, 8.0, 48, 500, 3),
    (-0.6, 70.0, 7.0, 30, 37, 1),
    (0.99, 88.0, 8.0, 30, 130, 2),
    (0.5, 140.0, 6.0, 24, 50, 1),
    (5.0e-5, 45.0, 8.0, 24, 80, 1),
    (0.0, 60.0, 8.0, 24, 120, 2),  # |a| < MINSPIN: theta walks at k = 0 (sn = sin, cn = cos)
    (-0.9, 80.0, 10.0, 40, 400, 7),
    (-0.3, 3.0, 12.0, 40, 30, 1),  # ngeo below one segment: queued ranges end at the last sample
    (0.082, 62.9, 8.0, 32, 160, 4),  # low spin: many radial blocks go through the repair pass
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"a{c[0]}_th{c[1]}_G{c[4]}_ch{c[5]}")
def test_walking_loops_match_direct_evaluation(case):
    a, thd, w, n, G, chunks = case
    A, B = grid(w, n)
    _walk_matches_direct(a, thd, A, B, G, chunks)


def test_case2_rays_with_all_three_integrals_transformed():
    """Radial case 2 (four real roots, the ray outside r4) has n1 > 1 and also
    n_+- = (r_+- - r3)(r4 - r1) / ((r_+- - r4)(r3 - r1)) > 1, as r_+- < r4.
    Just outside the shadow's prograde edge at high inclination r4 nears r_+
    and the transform (pt >= 0.05) takes I_+ and I_- too, so a wave walks all
    three log terms there rather than I_1's alone.  The screen below holds
    such rays next to rays that transform I_1 only."""
    a, thd = 0.99, 100.0
    al, be = np.linspace(-1.9, -1.5, 20), np.linspace(-0.7, 0.7, 24)
    A, B = (x.flatten() for x in np.meshgrid(al, be))
    s = Host().setup(a, 1.0e4, np.deg2rad(thd), A, B)
    r1, r2, r3, r4 = s["roots"][:, 0::2].T
    m = ((r3 - r2) * (r4 - r1)) / ((r3 - r1) * (r4 - r2))
    rp = 1.0 + np.sqrt(1.0 - a * a)
    with np.errstate(all="ignore"):
        n = ((rp - r3) * (r4 - r1)) / ((rp - r4) * (r3 - r1))
        pt = np.sqrt((n - 1.0) * (1.0 - m / n))
    hit = (s["rcase"] == 2) & (m >= 0.0) & (m < 1.0) & (n > 1.0) & (pt >= 0.05)
    assert 10 <= int(hit.sum()) < len(A) // 2, int(hit.sum())
    _walk_matches_direct(a, thd, A, B, 120, 2)


def _walk_matches_direct(a, thd, A, B, G, chunks):
    A, B = np.ascontiguousarray(A), np.ascontiguousarray(B)
    th = np.deg2rad(thd)
    host = Host()
    walk = host.raytrace_loops(a, 1.0e4, th, A, B, G, chunks=chunks)
    direct = host.raytrace_loops(a, 1.0e4, th, A, B, G, tau_in=walk["tau"], chunks=chunks)
    # samples where the direct evaluation itself moves by > 1e-12 under a
    # 1e-14 relative change of tau (log poles of the third-kind integrals,
    # the end of horizon-bound rays) are compared at that looser level
    nudged = host.raytrace_loops(a, 1.0e4, th, A, B, G, tau_in=walk["tau"] * (1.0 + 1.0e-14), chunks=chunks)
    assert np.array_equal(walk["status"], direct["status"])
    for k in ("r", "th", "ph", "t", "sig"):
        assert np.array_equal(np.isnan(walk[k]), np.isnan(direct[k])), k
        e = np.nan_to_num(errors(walk[k], direct[k], k))
        sens = np.nan_to_num(errors(nudged[k], direct[k], k))
        ill = sens > 1e-12
        assert ill.mean() < 0.05, (k, float(ill.mean()))
        # the walk adds at most ~1e-11 (normwise) to the direct evaluation:
        # an order of magnitude inside the 1e-10 parity bar
        assert e[~ill].max() < 2e-11, (k, float(e[~ill].max()))
        assert np.all(e[ill] < 100 * sens[ill] + 2e-11), k


def test_repair_pass_runs():
    """Radial long steps across a pole of a third-kind integral hand their
    blocks to the repair pass (the accuracy of those samples is checked with
    the rest in test_walking_loops_match_direct_evaluation)."""
    host = Host()
    for a, thd, G in ((0.082, 62.9, 160), (0.94, 17.0, 500)):
        A, B = grid(8.0, 32)
        out = host.raytrace_loops(a, 1.0e4, np.deg2rad(thd), np.ascontiguousarray(A), np.ascontiguousarray(B), G,
                                  chunks=4)
        assert out["nfix"][2] > 0, (a, thd)


PASS_CASES = [
    (0.94, 17.0, 8.0, 40, 500, 1),
    (0.94, 17.0, 8.0, 40, 500, 3),
    (-0.6, 70.0, 7.0, 30, 37, 1),
    (0.5, 140.0, 6.0, 24, 50, 2),
    (5.0e-5, 45.0, 8.0, 24, 80, 1),
    (0.3, 60.0, 5.0, 12, 4, 1),
]


@pytest.mark.parametrize("case", PASS_CASES, ids=lambda c: f"a{c[0]}_th{c[1]}_G{c[4]}_ch{c[5]}")
def test_three_passes_match_one_pass(case):
    """The step / anchor / walk / fixup passes (the kernels' path on the
    linspace grid) against one combined loop per coordinate.  Both walk the
    same grid steps; they differ in how a segment's first sample is reached
    (a long addition step from the last direct evaluation of the block, or a
    direct evaluation), so they agree to the walk's own accuracy: 4e-11 on
    samples the direct evaluation resolves (as in the test above), with the
    same status bits; the fixup queue is exercised."""
    a, thd, w, n, G, chunks = case
    A, B = grid(w, n)
    th = np.deg2rad(thd)
    host = Host()
    three = host.raytrace_loops(a, 1.0e4, th, A, B, G, chunks=chunks, passes=3)
    one = host.raytrace_loops(a, 1.0e4, th, A, B, G, chunks=chunks, passes=1)
    direct = host.raytrace_loops(a, 1.0e4, th, A, B, G, tau_in=three["tau"], chunks=chunks)
    nudged = host.raytrace_loops(a, 1.0e4, th, A, B, G, tau_in=three["tau"] * (1.0 + 1.0e-14), chunks=chunks)
    assert np.array_equal(three["status"], one["status"])
    assert three["nfix"].sum() > 0
    for k in ("tau", "t", "r", "th", "ph", "sig"):
        assert np.array_equal(np.isnan(three[k]), np.isnan(one[k])), k
        e = np.nan_to_num(errors(three[k], one[k], k))
        ill = np.nan_to_num(errors(nudged[k], direct[k], k)) > 1e-12
        assert e[~ill].max() < 4e-11, (k, float(e[~ill].max()))


def test_short_grid_queued_ranges_fit_the_queue():
    """ngeo = 7 is shorter than one segment: a refused step queues the samples
    up to the last one only, so the kernels' queue (ceil(ngeo/8) items per
    pixel and pass) holds every range (raytrace_loops asserts it), and the
    walked samples match direct evaluation."""
    A, B = grid(12.0, 40)
    th = np.deg2rad(3.0)
    host = Host()
    walk = host.raytrace_loops(-0.3, 1.0e4, th, A, B, 7, passes=3)
    direct = host.raytrace_loops(-0.3, 1.0e4, th, A, B, 7, tau_in=walk["tau"])
    assert walk["nfix"].sum() > 0
    for k in ("r", "th", "ph", "t", "sig"):
        assert np.array_equal(np.isnan(walk[k]), np.isnan(direct[k])), k
        e = np.nan_to_num(errors(walk[k][:-1], direct[k][:-1], k))
        assert e.max() < 2e-11, (k, float(e.max()))
