"""CPU oracle for the analytic Kerr raytracer.

TEST INFRASTRUCTURE ONLY.  Nothing on the product path imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may use it, and only as the checker / the timed CPU
baseline, never as the thing measured or shipped.

What it is
----------
A NumPy/SciPy restatement of the reference's algorithm for the hot path
(achael/kgeo, Gralla & Lupsasca 2019a,b formalism).  Every function names
the reference file:line it follows.  The restatement keeps the reference's
arithmetic: the same scipy.special calls (Cephes ``ellipkinc``/``ellipeinc``/
``ellipj``/``ellipk`` and Carlson ``elliprf``/``elliprj``/``elliprc``), the same
numpy complex semantics in the root finder (Smith division, ``th*(1/3)``), the
same masks, clamps and special cases, so that its outputs are the reference's
outputs up to the reference's own rounding.  It is vectorised over
``(ngeo, npix)`` arrays exactly as the reference is, so its timing is a fair
stand-in for the reference CPU path.

How it is pinned (tests/test_oracle_*.py)
-----------------------------------------
* Elliptic layer: mpmath ``ellippi``/``ellipf``/``ellipe`` at 30 digits
  (principal value for n>1), periodicity and sign symmetry, and every
  special-case branch of ``ellip_pi`` (NaN / inf / limits).
* Roots: residual of the radial potential R(r), Vieta's relations, and
  mpmath ``polyroots`` at 50 digits for the labels.
* Mino time: ``mpmath.quad`` of dr/sqrt(R) for every radial class, region
  II (rays from inside the outer horizon's turning region) included.
* Geodesic samples: initial conditions at tau=0, the Mino-time geodesic ODE
  residual (first integrals (dr/dtau)^2=R, (dth/dtau)^2=Theta, dphi/dtau and
  dt/dtau by high-order finite differences), the equatorial consistency of
  r_equatorial vs coords_at_tau (also south of the equator), and the 15
  ipole midplane-crossing vectors that the reference's own test holds.
* Counts: equatorial crossings against sign changes of cos(theta) at every
  sample, for observers north and south of the equator; poloidal orbits
  against one period found by root-finding on the sampled ray; angular
  turning points (ordinary and vortical) against the extrema of theta; the
  radial turning count against the extrema of r.
* Redshift model: u.u = -1 for the Keplerian/plunging velocity; g against
  -1/(k.u) with k from finite differences of the traced ray; the
  Schwarzschild circular-orbit closed form.
* Lensed curves: returned radii are sign changes of r_equatorial2 - r_eq;
  Schwarzschild curves are mirror-symmetric; critical-curve points are double
  roots of R(r) and separate rays inside from rays outside.

Parity unpinned (stated here and in the design notes): ``my_cbrt`` on an exact tie of
real parts (the reference returns uninitialised memory there; this
restatement returns NaN), and the reference's printed warnings.
"""

import numpy as np
import scipy.special as sp

MINSPIN = 1.0e-4        # kerr_raytracing_utils.py:9
EP = 1.0e-12            # kerr_raytracing_utils.py:12
MAXTAUFRAC = 1.0 - 1.0e-10   # kerr_raytracing_ana.py:26
HALFPI = 0.5 * np.pi    # scipy_ellip_binding.py:7


# ---------------------------------------------------------------------------
# small helpers
# ---------------------------------------------------------------------------

def _as1d(x):
    if not isinstance(x, np.ndarray):
        x = np.array([x]).flatten()
    return x


def check_params(a, r_o=None, th_o=None):
    """Argument validation of kerr_raytracing_ana.py:58-66 (same messages)."""
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")
    if r_o is not None:
        if not (isinstance(r_o, float) and (r_o >= 100)):
            raise Exception("r_o should be a float > 100")
    if th_o is not None:
        if not (isinstance(th_o, float) and (0 < th_o < np.pi) and th_o != 0.5 * np.pi):
            raise Exception("th_o should be a float in range (0,pi/2) or (pi/2,pi)")


def sign_ge0(x):
    """+1 for x>=0, -1 for x<0, 0 for NaN (kerr_raytracing_utils.py:857-864)."""
    out = np.zeros(x.shape)
    out[x >= 0] = 1
    out[x < 0] = -1
    return out


def conserved(a, th_o, alpha, beta):
    """lambda, eta from screen coordinates (kerr_raytracing_ana.py:76-77)."""
    lam = -alpha * np.sin(th_o)
    eta = (alpha ** 2 - a ** 2) * np.cos(th_o) ** 2 + beta ** 2
    return lam, eta


def conserved_fixed(a, th_o, alpha, beta):
    """conserved() plus the spin-zero vortical fix (kerr_raytracing_ana.py:80-82)."""
    lam, eta = conserved(a, th_o, alpha, beta)
    if np.abs(a) < MINSPIN and np.any(eta < 0):
        eta[eta < 0] = EP
    return lam, eta


# ---------------------------------------------------------------------------
# angular turning points  (kerr_raytracing_utils.py:242-328)
# ---------------------------------------------------------------------------

def uplus_uminus(a, th_o, lam, eta):
    """u_+, u_-, u_+/u_-, a^2 u_- with the reference's clamps.

    kerr_raytracing_utils.py:276-328.  For |a|<MINSPIN the reference binds
    u_minus to the *same array* as u_plus (:291-292), so every later in-place
    clamp applies to both; this is reproduced by aliasing here too.
    """
    c2 = np.cos(th_o) ** 2
    if np.abs(a) < MINSPIN:
        up = eta / (eta + lam * lam)
        um = up
    else:
        dth = 0.5 * (1 - (eta + lam * lam) / (a * a))
        disc = np.sqrt(dth ** 2 + eta / (a * a))
        up = dth + disc
        um = dth - disc
    up[(c2 - up) > 0] = c2
    up[up ** 2 > 1] = 1
    um[(um == 0) * (eta < 0)] = EP
    um[(um == 0) * (eta >= 0)] = -EP
    vort = eta < 0
    clamp = np.zeros(eta.shape, dtype=bool)
    clamp[vort] = (c2 - um[vort]) < 0
    um[clamp] = c2
    if np.abs(a) < MINSPIN:
        ratio = 0. * um
        a2um = -(eta + lam ** 2)
    else:
        ratio = up / um
        a2um = a ** 2 * um
    return up, um, ratio, a2um


def theta_class(eta):
    """thclass: 1 eta>0, 2 eta<0, 3 eta==0, 0 NaN (kerr_raytracing_utils.py:266-272)."""
    cls = np.zeros(eta.shape, dtype=int)
    cls[eta > 0] = 1
    cls[eta < 0] = 2
    cls[eta == 0] = 3
    return cls


# ---------------------------------------------------------------------------
# radial roots  (kerr_raytracing_utils.py:330-380, 828-855)
# ---------------------------------------------------------------------------

def principal_cbrt(x):
    """Cube root of a complex array as kerr_raytracing_utils.py:828-855.

    Real inputs (imag == 0) take sign(x)*cbrt|x|; complex inputs take the
    root with the largest real part, built from atan2 and complex exp with
    numpy's complex semantics.  Exact ties of real parts (which the reference
    leaves as uninitialised memory) are returned as NaN here.
    """
    sgn = np.sign(x)
    mag = np.abs(x)
    ang = np.angle(x)
    out = np.full(x.shape, np.nan + 0j)
    real = np.imag(x) == 0
    out[real] = (sgn * np.cbrt(mag))[real]
    if np.any(~real):
        c1 = np.cbrt(mag) * np.exp(1j * ang / 3.)
        c2 = np.cbrt(mag) * np.exp(1j * ang / 3. + 2j * np.pi / 3.)
        c3 = np.cbrt(mag) * np.exp(1j * ang / 3. - 2j * np.pi / 3.)
        x1, x2, x3 = np.real(c1), np.real(c2), np.real(c3)
        pick1 = ~real * (x1 > x2) * (x1 > x3)
        pick2 = ~real * (x2 > x1) * (x2 > x3)
        pick3 = ~real * (x3 > x1) * (x3 > x2)
        out[pick1] = c1[pick1]
        out[pick2] = c2[pick2]
        out[pick3] = c3[pick3]
    return out


def radial_roots(a, lam, eta):
    """r1..r4 (complex) and rclass (kerr_raytracing_utils.py:330-380)."""
    rplus = 1 + np.sqrt(1 - a ** 2)
    A = a ** 2 - eta - lam ** 2
    B = 2 * (eta + (lam - a) ** 2)
    C = -(a ** 2) * eta
    P = -(A ** 2) / 12. - C
    Q = (-A / 3.) * ((A ** 2) / 36. - C) - (B ** 2) / 8.
    disc = (4 * (P ** 3) + 27 * (Q ** 2)).astype(complex)
    wp3 = -0.5 * Q + np.sqrt(disc / 108.)
    wm3 = -0.5 * Q - np.sqrt(disc / 108.)
    xi0 = principal_cbrt(wp3) + principal_cbrt(wm3) - A / 3.
    z = np.sqrt(0.5 * xi0)
    w_lo = (-0.5 * A - z ** 2 + 0.25 * B / z).astype(complex)
    w_hi = (-0.5 * A - z ** 2 - 0.25 * B / z).astype(complex)
    r1 = -z - np.sqrt(w_lo)
    r2 = -z + np.sqrt(w_lo)
    r3 = z - np.sqrt(w_hi)
    r4 = z + np.sqrt(w_hi)
    rclass = np.zeros(r1.shape, dtype=int)
    im2 = np.imag(r2) != 0
    im4 = np.imag(r4) != 0
    rclass[im2] = 4
    rclass[~im2 * im4] = 3
    rclass[~im2 * ~im4 * (np.real(r4) < rplus)] = 2
    rclass[~im2 * ~im4 * (np.real(r4) > rplus)] = 1
    return r1, r2, r3, r4, rclass


# ---------------------------------------------------------------------------
# elliptic integral of the third kind  (scipy_ellip_binding.py:87-237)
# ---------------------------------------------------------------------------

def ellip_pi(n, phi, m):
    """Pi(n; phi | m), A&S convention, as scipy_ellip_binding.py:87-237.

    Special cases in the reference's order (:125-162), periodic reduction
    (:164-192) and the Carlson form with the DLMF 19.20.14 transformation for
    rho <= 0 (:194-235), which yields the Cauchy principal value for n > 1.
    """
    n = _as1d(n)
    phi = _as1d(phi)
    m = _as1d(m)
    if not (n.shape[-1] == phi.shape[-1] == m.shape[-1]):
        raise Exception("inputs to ellip_pi_arr do not have same last dimension!")
    if len(n.shape) != len(phi.shape):
        n = np.outer(np.ones(phi.shape[0]), n)
    if len(m.shape) != len(phi.shape):
        m = np.outer(np.ones(phi.shape[0]), m)

    cphi = np.cos(phi)
    sphi = np.sin(phi)
    x = cphi ** 2
    y = 1. - m * sphi ** 2
    z = np.ones(x.shape)
    rho = 1. - n * sphi ** 2

    out = np.zeros(phi.shape)
    done = np.zeros(phi.shape, dtype=bool)

    def take(mask, value):
        nonlocal done
        if np.any(mask):
            out[mask] = value if np.isscalar(value) else value[mask]
        done = done + mask

    take(np.abs(phi) <= 1.e-10, phi)
    take((m < -1.e14) * ~done, 0.)
    take((y < 0) * ~done, np.nan)
    take((np.abs(phi) > HALFPI) * (m >= 1) * ~done, np.nan)
    take((np.abs(phi) > HALFPI) * (n == 1) * ~done, np.inf)
    take((rho == 0) * ~done, np.inf)
    take(np.isinf(n) * ~done, 0.)

    # several periods: Pi(phi) = 2k Pi(n|m) + Pi(phi - k pi)
    big = (np.abs(phi) > np.pi) * ~done
    if np.any(big):
        mb, nb, pb = m[big], n[big], phi[big]
        sg = np.sign(pb)
        kk = sg * np.floor(np.abs(pb) / np.pi)
        red = pb - kk * np.pi
        out[big] = 2 * kk * ellip_pi(nb, HALFPI * np.ones(red.shape), mb) + ellip_pi(nb, red, mb)
    done = done + big

    # one half period: Pi(phi) = 2 s Pi(n|m) - Pi(s pi - phi)
    half = (np.abs(phi) > HALFPI) * (np.abs(phi) <= np.pi) * ~done
    if np.any(half):
        mh, nh, ph = m[half], n[half], phi[half]
        sg = np.sign(ph)
        red = -ph + sg * np.pi
        out[half] = sg * 2 * ellip_pi(nh, HALFPI * np.ones(red.shape), mh) + (-ellip_pi(nh, red, mh))
    done = done + half

    rest = ~done
    if np.any(rest):
        nr = n[rest]
        sr = sphi[rest]
        xr, yr, zr, rr = x[rest], y[rest], z[rest], rho[rest]
        qr = -rr
        pr = (zr * (xr + yr + qr) - xr * yr) / (zr + qr)
        rf = sp.elliprf(xr, yr, zr)
        rj = np.zeros(nr.shape)
        pos = rr > 0
        if np.any(pos):
            rj[pos] = sp.elliprj(xr[pos], yr[pos], zr[pos], rr[pos])
        neg = ~pos
        if np.any(neg):
            # DLMF 19.20.14: R_J(x,y,z,-q) from R_J(x,y,z,p) and R_C
            rj0 = sp.elliprj(xr[neg], yr[neg], zr[neg], pr[neg])
            rc = sp.elliprc((xr * yr + pr * qr)[neg], (pr * qr)[neg])
            fac = 3 * np.sqrt((xr * yr * zr)[neg] / (xr * yr + pr * qr)[neg])
            rj[neg] = ((pr - zr)[neg] * rj0 - 3 * rf[neg] + fac * rc)
            rj[neg] = rj[neg] / (qr + zr)[neg]
        out[rest] = sr * rf + (nr / 3.) * (sr ** 3) * rj
    return out


# ---------------------------------------------------------------------------
# Mino time to the end of the ray  (kerr_raytracing_utils.py:382-548)
# ---------------------------------------------------------------------------

def mino_total(a, r_o, eta, r1, r2, r3, r4):
    """tau_tot per pixel (kerr_raytracing_utils.py:382-548).

    Region IV/III/II rays end on the horizon, region I rays return to
    infinity after one radial turning point.  Regions II and I are split on
    r3 vs r_+ (:485,518), not on r4 as rclass is.
    """
    rplus = 1 + np.sqrt(1 - a ** 2)
    out = np.zeros(r1.shape)

    m4 = np.imag(r1) != 0
    if np.any(m4):
        a1, b1 = np.imag(r4)[m4], np.real(r4)[m4]
        a2, b2 = np.imag(r2)[m4], np.real(r2)[m4]
        CC = np.sqrt((a1 - a2) ** 2 + (b1 - b2) ** 2)
        DD = np.sqrt((a1 + a2) ** 2 + (b1 - b2) ** 2)
        k4 = (4 * CC * DD) / ((CC + DD) ** 2)
        g0 = np.sqrt((4 * a2 ** 2 - (CC - DD) ** 2) / ((CC + DD) ** 2 - 4 * a2 ** 2))
        x4rp = (rplus + b1) / a2
        x4ro = np.inf if r_o == np.inf else (r_o + b1) / a2
        pref = 2. / (CC + DD)
        I_hor = pref * sp.ellipkinc(np.arctan(x4rp) + np.arctan(g0), k4)
        I_obs = pref * sp.ellipkinc(np.arctan(x4ro) + np.arctan(g0), k4)
        out[m4] = I_obs - I_hor

    m3 = (np.imag(r3) != 0) * (~m4)
    if np.any(m3):
        a1, b1 = np.imag(r4)[m3], np.real(r4)[m3]
        rr1, rr2 = np.real(r1)[m3], np.real(r2)[m3]
        rr21 = rr2 - rr1
        rrp1, rrp2 = rplus - rr1, rplus - rr2
        rro1, rro2 = r_o - rr1, r_o - rr2
        AA = np.sqrt(a1 ** 2 + (b1 - rr2) ** 2)
        BB = np.sqrt(a1 ** 2 + (b1 - rr1) ** 2)
        k3 = ((AA + BB) ** 2 - rr21 ** 2) / (4 * AA * BB)
        x3rp = (AA * rrp1 - BB * rrp2) / (AA * rrp1 + BB * rrp2)
        if r_o == np.inf:
            x3ro = (AA - BB) / (AA + BB)
        else:
            x3ro = (AA * rro1 - BB * rro2) / (AA * rro1 + BB * rro2)
        pref = 1. / np.sqrt(AA * BB)
        I_hor = pref * sp.ellipkinc(np.arccos(x3rp), k3)
        I_obs = pref * sp.ellipkinc(np.arccos(x3ro), k3)
        out[m3] = I_obs - I_hor

    m2 = (np.imag(r3) == 0.) * (r3 < rplus) * (~m3)
    if np.any(m2):
        rr1, rr2, rr3, rr4 = (np.real(r)[m2] for r in (r1, r2, r3, r4))
        rr31, rr32, rr41, rr42 = rr3 - rr1, rr3 - rr2, rr4 - rr1, rr4 - rr2
        k2 = (rr32 * rr41) / (rr31 * rr42)
        x2rp = np.sqrt((rr31 * (rplus - rr4)) / (rr41 * (rplus - rr3)))
        if r_o == np.inf:
            x2ro = np.sqrt(rr31 / rr41)
        else:
            x2ro = np.sqrt((rr31 * (r_o - rr4)) / (rr41 * (r_o - rr3)))
        pref = 2. / np.sqrt(rr31 * rr42)
        I_hor = pref * sp.ellipkinc(np.arcsin(x2rp), k2)
        I_obs = pref * sp.ellipkinc(np.arcsin(x2ro), k2)
        out[m2] = I_obs - I_hor

    m1 = (np.imag(r3) == 0.) * (r3 > rplus) * (~m2)
    if np.any(m1):
        rr1, rr2, rr3, rr4 = (np.real(r)[m1] for r in (r1, r2, r3, r4))
        rr31, rr32, rr41, rr42 = rr3 - rr1, rr3 - rr2, rr4 - rr1, rr4 - rr2
        k2 = (rr32 * rr41) / (rr31 * rr42)
        if r_o == np.inf:
            x2ro = np.sqrt(rr31 / rr41)
        else:
            x2ro = np.sqrt((rr31 * (r_o - rr4)) / (rr41 * (r_o - rr3)))
        pref = 2. / np.sqrt(rr31 * rr42)
        I_obs = pref * sp.ellipkinc(np.arcsin(x2ro), k2)
        out[m1] = 2 * I_obs
    return out


# ---------------------------------------------------------------------------
# theta integrals  (kerr_raytracing_ana.py:195-365)
# ---------------------------------------------------------------------------

def _amplitude_negative_m(snarg, k):
    """sin(am(x|k)) and am(x|k) for k<0 via the imaginary-modulus identity
    with m' = k/(k-1) (kerr_raytracing_ana.py:262-277, 337-352)."""
    s_out = np.zeros(snarg.shape)
    am_out = np.zeros(snarg.shape)
    tiny = np.abs(snarg) < EP
    if np.any(tiny):
        s_out[tiny] = snarg[tiny]
        am_out[tiny] = snarg[tiny]
    if np.any(~tiny):
        mp = k / (k - 1)
        mp = np.outer(np.ones(snarg.shape[0]), mp)[~tiny]
        xs = snarg[~tiny] / np.sqrt(1 - mp)
        sn, cn, dn, _ = sp.ellipj(xs, mp)
        s_out[~tiny] = np.sqrt(1 - mp) * sn / dn
        am_out[~tiny] = 0.5 * np.pi - sp.ellipj(sp.ellipk(mp) - xs, mp)[3]
    return s_out, am_out


def th_integrate(a, th_o, s_o, lam, eta, u_plus, u_minus, tau, do_phi_and_t=True):
    """theta(tau), G_phi(tau), G_t(tau) (kerr_raytracing_ana.py:195-365)."""
    th_s = np.zeros(tau.shape)
    G_ph = np.zeros(tau.shape)
    G_t = np.zeros(tau.shape)

    # ordinary motion, eta >= 0 (taken only if some eta > 0, :217-218)
    if np.any(eta > 0.):
        msk = eta >= 0.
        up, um, s = u_plus[msk], u_minus[msk], s_o[msk]
        if np.abs(a) < MINSPIN:
            k = np.zeros(eta.shape)[msk]
            a2um = -(eta + lam ** 2)[msk]
        else:
            k = up / um
            a2um = a ** 2 * um
        pref = 1 / np.sqrt(-a2um)
        arg_o = np.arcsin(np.cos(th_o) / np.sqrt(up))
        F_o = sp.ellipkinc(arg_o, k)
        Gth_o = -pref * F_o
        if do_phi_and_t:
            Gph_o = -pref * ellip_pi(up, arg_o, k)
            k0 = k == 0
            Gt_o = np.zeros(k.shape)
            if np.any(k0):
                Gt_o[k0] = 2 * up * pref * 0.125 * (-2 * arg_o + np.sin(2 * arg_o))[k0]
            if np.any(~k0):
                Gt_o[~k0] = 2 * up * pref * (sp.ellipeinc(arg_o, k) - F_o)[~k0] / (2 * k[~k0])
        snarg = (np.sqrt(-a2um) * (-tau[:, msk] + s * Gth_o)).astype(float)
        sinPhi, Phi = _amplitude_negative_m(snarg, k)
        th_s[:, msk] = (np.arccos(-s * np.sqrt(up) * sinPhi)).astype(float)
        if do_phi_and_t:
            G_ph[:, msk] = (pref * ellip_pi(up, Phi, k) - s * Gph_o).astype(float)
            k0 = k == 0
            gt = np.zeros(Phi.shape)
            if np.any(k0):
                gt[:, k0] = -2 * up * pref * 0.125 * (-2 * Phi + np.sin(2 * Phi))[:, k0]
            if np.any(~k0):
                gt[:, ~k0] = -(2 * up * pref * (sp.ellipeinc(Phi, k) - sp.ellipkinc(Phi, k)))[:, ~k0] / (2 * k[~k0])
            G_t[:, msk] = gt - s * Gt_o

    # vortical motion, eta < 0 (:300-363)
    if np.any(eta <= 0.):
        if np.abs(a) < MINSPIN:
            raise Exception("below MINSPIN but there are eta<0 points in th_integrate!")
        msk = eta < 0.
        up, um, s = u_plus[msk], u_minus[msk], s_o[msk]
        a2um = a ** 2 * um
        h = 1. if th_o <= np.pi / 2 else -1
        prefA = 1 / np.sqrt(a2um)
        prefB = prefA / (1. - um)
        prefC = np.sqrt(um / a ** 2)
        Nu_o = np.arcsin(np.sqrt((np.cos(th_o) ** 2 - um) / (up - um)))
        k = 1 - up / um
        upm = (up - um) / (1. - um)
        Gth_o = -h * prefA * sp.ellipkinc(Nu_o, k)
        if do_phi_and_t:
            Gph_o = -h * prefB * ellip_pi(upm, Nu_o, k)
            Gt_o = -h * prefC * sp.ellipeinc(Nu_o, k)
        snarg = (np.sqrt(a2um) * (-tau[:, msk] + s * Gth_o)).astype(float)
        sinNu, Nu = _amplitude_negative_m(snarg, k)
        th_s[:, msk] = (np.arccos(h * np.sqrt(um + (up - um) * sinNu ** 2))).astype(float)
        if do_phi_and_t:
            G_ph[:, msk] = (prefB * ellip_pi(upm, Nu, k) - s * Gph_o).astype(float)
            G_t[:, msk] = (prefC * sp.ellipeinc(Nu, k) - s * Gt_o).astype(float)
    return th_s, G_ph, G_t


# ---------------------------------------------------------------------------
# radial integrals  (kerr_raytracing_ana.py:368-683)
# ---------------------------------------------------------------------------

def S1_S2(al, phi, j, ret_s2=True):
    """GL19a B92/B93 building blocks (kerr_raytracing_ana.py:638-663)."""
    al2 = al * al
    al2p1 = 1 + al2
    al2j = 1 - j + al2
    s2phi = np.sqrt(1. - j * (np.sin(phi)) ** 2)
    p2 = np.sqrt(al2p1 / al2j)
    f2 = 0.5 * p2 * np.log(np.abs(((1 - p2) * (1 + p2 * s2phi)) / ((1 + p2) * (1 - p2 * s2phi))))
    F = sp.ellipkinc(phi, j)
    S1 = (F + al2 * ellip_pi(al2p1, phi, j) - al * f2) / al2p1
    S2 = np.nan
    if ret_s2:
        E = sp.ellipeinc(phi, j)
        tphi = np.tan(phi)
        S2a = (1 - j) * F + al2 * E + al2 * s2phi * (al - tphi) / (1 + al * tphi) - al2 * al
        S2a = -S2a / (al2p1 * al2j)
        S2b = S1 * (1. / al2p1 + (1 - j) / al2j)
        S2 = S2a + S2b
    return S1, S2


def R1_R2(al, phi, j, ret_r2=True):
    """GL19a B62/B65 building blocks (kerr_raytracing_ana.py:666-683)."""
    al2 = al ** 2
    s2phi = np.sqrt(1 - j * np.sin(phi) ** 2)
    p1 = np.sqrt((al2 - 1) / (j + (1 - j) * al2))
    f1 = 0.5 * p1 * np.log(np.abs((p1 * s2phi + np.sin(phi)) / (p1 * s2phi - np.sin(phi))))
    nn = al2 / (al2 - 1)
    R1 = (ellip_pi(nn, phi, j) - al * f1) / (1 - al2)
    R2 = np.nan
    if ret_r2:
        F = sp.ellipkinc(phi, j)
        E = sp.ellipeinc(phi, j)
        R2 = (F - (al2 / (j + (1 - j) * al2)) * (E - al * np.sin(phi) * s2phi / (1 + al * np.cos(phi)))) / (al2 - 1)
        R2 = R2 + (2 * j - nn) * R1 / (j + (1 - j) * al2)
    return R1, R2


def r_integrate(a, r_o, lam, eta, r1, r2, r3, r4, tau, do_phi_and_t=True):
    """r(tau), I_phi, I_t, I_sigma (kerr_raytracing_ana.py:368-633)."""
    rplus = 1 + np.sqrt(1 - a ** 2)
    rminus = 1 - np.sqrt(1 - a ** 2)
    r_s = np.zeros(tau.shape)
    I_p = np.zeros(tau.shape)
    I_m = np.zeros(tau.shape)
    I_1 = np.zeros(tau.shape)
    I_2 = np.zeros(tau.shape)
    s = 1

    # case 4: no real roots (:404-471)
    m4 = np.imag(r1) != 0
    if np.any(m4):
        tt = tau[:, m4]
        a1, b1 = np.imag(r4)[m4], np.real(r4)[m4]
        a2, b2 = np.imag(r2)[m4], np.real(r2)[m4]
        CC = np.sqrt((a1 - a2) ** 2 + (b1 - b2) ** 2)
        DD = np.sqrt((a1 + a2) ** 2 + (b1 - b2) ** 2)
        k4 = (4 * CC * DD) / ((CC + DD) ** 2)
        x4rp = (rplus + b1) / a2
        x4rm = (rminus + b1) / a2
        x4ro = np.inf if r_o == np.inf else (r_o + b1) / a2
        g0 = np.sqrt((4 * a2 ** 2 - (CC - DD) ** 2) / ((CC + DD) ** 2 - 4 * a2 ** 2))
        gp = (g0 * x4rp - 1.) / (g0 + x4rp)
        gm = (g0 * x4rm - 1.) / (g0 + x4rm)
        aux = np.arctan(x4ro) + np.arctan(g0)
        pref = 2. / (CC + DD)
        Ir_o = pref * sp.ellipkinc(aux, k4)
        X4 = (1. / pref) * (s * (-tt) + Ir_o)
        sn, cn, dn, am = sp.ellipj(X4, k4)
        sc = sn / cn
        r_s[:, m4] = -a2 * (g0 - sc) / (1 + g0 * sc) - b1
        if do_phi_and_t:
            S1a0, S2a0 = S1_S2(g0, am, k4, ret_s2=True)
            S1b0, S2b0 = S1_S2(g0, aux, k4, ret_s2=True)
            S1ap, _ = S1_S2(gp, am, k4, ret_s2=False)
            S1bp, _ = S1_S2(gp, aux, k4, ret_s2=False)
            S1am, _ = S1_S2(gm, am, k4, ret_s2=False)
            S1bm, _ = S1_S2(gm, aux, k4, ret_s2=False)
            aa0 = (a2 / g0) * (1 + g0 ** 2)
            Pi1 = s * pref * aa0 * (S1a0 - S1b0)
            Pi2 = s * pref * (aa0 ** 2) * (S2a0 - S2b0)
            Pip = s * pref * (1 + g0 ** 2) / (g0 ** 2 + g0 * x4rp) * (S1ap - S1bp)
            Pim = s * pref * (1 + g0 ** 2) / (g0 ** 2 + g0 * x4rm) * (S1am - S1bm)
            aa1 = (a2 / g0 - b1)
            I_1[:, m4] = aa1 * (-tt) - Pi1
            I_2[:, m4] = (aa1 ** 2) * (-tt) - 2 * aa1 * Pi1 + Pi2
            I_p[:, m4] = g0 / (a2 - a2 * g0 * x4rp) * (-tt - Pip)
            I_m[:, m4] = g0 / (a2 - a2 * g0 * x4rm) * (-tt - Pim)

    # case 3: r3, r4 complex (:473-552)
    m3 = (np.imag(r3) != 0) * (~m4)
    if np.any(m3):
        tt = tau[:, m3]
        a1, b1 = np.imag(r4)[m3], np.real(r4)[m3]
        rr1, rr2 = np.real(r1)[m3], np.real(r2)[m3]
        rr21 = rr2 - rr1
        rrp1, rrp2 = rplus - rr1, rplus - rr2
        rrm1, rrm2 = rminus - rr1, rminus - rr2
        rro1, rro2 = r_o - rr1, r_o - rr2
        AA = np.sqrt(a1 ** 2 + (b1 - rr2) ** 2)
        BB = np.sqrt(a1 ** 2 + (b1 - rr1) ** 2)
        k3 = ((AA + BB) ** 2 - rr21 ** 2) / (4 * AA * BB)
        x3rp = (AA * rrp1 - BB * rrp2) / (AA * rrp1 + BB * rrp2)
        x3rm = (AA * rrm1 - BB * rrm2) / (AA * rrm1 + BB * rrm2)
        if r_o == np.inf:
            x3ro = (AA - BB) / (AA + BB)
        else:
            x3ro = (AA * rro1 - BB * rro2) / (AA * rro1 + BB * rro2)
        alp = -1. / x3rp
        alm = -1. / x3rm
        al0 = (BB + AA) / (BB - AA)
        aux = np.arccos(x3ro)
        pref = (1. / np.sqrt(AA * BB))
        Ir_o = pref * sp.ellipkinc(aux, k3)
        X3 = (1. / pref) * (-tt + s * Ir_o)
        sn, cn, dn, am = sp.ellipj(X3, k3)
        num = (BB * rr2 - AA * rr1) + (BB * rr2 + AA * rr1) * cn
        den = (BB - AA) + (BB + AA) * cn
        r_s[:, m3] = num / den
        if do_phi_and_t:
            R1a0, R2a0 = R1_R2(al0, am, k3)
            R1b0, R2b0 = R1_R2(al0, aux, k3)
            R1ap, _ = R1_R2(alp, am, k3, ret_r2=False)
            R1bp, _ = R1_R2(alp, aux, k3, ret_r2=False)
            if np.abs(a) > MINSPIN:
                R1am, _ = R1_R2(alm, am, k3, ret_r2=False)
                R1bm, _ = R1_R2(alm, aux, k3, ret_r2=False)
            else:
                R1am = np.zeros(R1ap.shape)
                R1bm = np.zeros(R1ap.shape)
            c0 = (2 * rr21 * np.sqrt(AA * BB)) / (BB ** 2 - AA ** 2)
            Pi1 = c0 * (R1a0 - s * R1b0)
            Pi2 = c0 ** 2 * (R2a0 - s * R2b0)
            Pip = ((2 * rr21 * np.sqrt(AA * BB)) / (BB * rrp2 - AA * rrp1)) * (R1ap - s * R1bp)
            Pim = ((2 * rr21 * np.sqrt(AA * BB)) / (BB * rrm2 - AA * rrm1)) * (R1am - s * R1bm)
            pr = ((BB * rr2 + AA * rr1) / (BB + AA))
            I_1[:, m3] = pr * (-tt) + Pi1
            I_2[:, m3] = pr ** 2 * (-tt) + 2 * pr * Pi1 + np.sqrt(AA * BB) * Pi2
            I_p[:, m3] = -((BB + AA) * (-tt) + Pip) / (BB * rrp2 + AA * rrp1)
            I_m[:, m3] = -((BB + AA) * (-tt) + Pim) / (BB * rrm2 + AA * rrm1)

    # case 2: four real roots, regions I and II (:554-624)
    m2 = (np.imag(r3) == 0.) * (~m3)
    if np.any(m2):
        tt = tau[:, m2]
        rr1, rr2, rr3, rr4 = (np.real(r)[m2] for r in (r1, r2, r3, r4))
        rr31, rr32, rr41, rr42, rr43 = rr3 - rr1, rr3 - rr2, rr4 - rr1, rr4 - rr2, rr4 - rr3
        rrp3, rrm3 = rplus - rr3, rminus - rr3
        rrp4, rrm4 = rplus - rr4, rminus - rr4
        k2 = (rr32 * rr41) / (rr31 * rr42)
        if r_o == np.inf:
            x2ro = np.sqrt(rr31 / rr41)
        else:
            x2ro = np.sqrt((rr31 * (r_o - rr4)) / (rr41 * (r_o - rr3)))
        aux = np.arcsin(x2ro)
        Ir_o = 2 * sp.ellipkinc(aux, k2) / np.sqrt(rr31 * rr42)
        X2 = 0.5 * np.sqrt(rr31 * rr42) * (-tt + s * Ir_o)
        sn, cn, dn, am = sp.ellipj(X2, k2)
        rs = (rr4 * rr31 - rr3 * rr41 * (sn) ** 2) / (rr31 - rr41 * (sn) ** 2)
        r_s[:, m2] = rs
        if do_phi_and_t:
            dX = -0.5 * np.sqrt(rr31 * rr42)
            dsn2 = 2 * sn * cn * dn * dX
            drs = rr31 * rr43 * rr41 * dsn2 / ((rr31 - rr41 * sn ** 2) ** 2)
            drs *= -1
            Rpot_o = (r_o - rr1) * (r_o - rr2) * (r_o - rr3) * (r_o - rr4)
            drs_o = s * np.sqrt(Rpot_o)
            H = drs / (rs - rr3) - drs_o / (r_o - rr3)
            E = np.sqrt(rr31 * rr42) * (sp.ellipeinc(am, k2) - s * sp.ellipeinc(aux, k2))
            c = 2. / np.sqrt(rr31 * rr42)
            n1 = rr41 / rr31
            npl = (rrp3 * rr41) / (rrp4 * rr31)
            nmi = (rrm3 * rr41) / (rrm4 * rr31)
            Pi1 = c * (ellip_pi(n1, am, k2) - s * ellip_pi(n1, aux, k2))
            Pip = c * (rr43 / (rrp3 * rrp4)) * (ellip_pi(npl, am, k2) - s * ellip_pi(npl, aux, k2))
            Pim = c * (rr43 / (rrm3 * rrm4)) * (ellip_pi(nmi, am, k2) - s * ellip_pi(nmi, aux, k2))
            I_1[:, m2] = rr3 * (-tt) + rr43 * Pi1
            I_2[:, m2] = H - 0.5 * (rr1 * rr4 + rr2 * rr3) * (-tt) - E
            I_p[:, m2] = tt / rrp3 - Pip
            I_m[:, m2] = tt / rrm3 - Pim

    I_0 = -tau
    I_phi = (2 * a / (rplus - rminus)) * ((rplus - 0.5 * a * lam) * I_p - (rminus - 0.5 * a * lam) * I_m)
    I_tA = (4 / (rplus - rminus)) * ((rplus ** 2 - 0.5 * a * lam * rplus) * I_p - (rminus ** 2 - 0.5 * a * lam * rminus) * I_m)
    I_t = I_tA + 4 * I_0 + 2 * I_1 + I_2
    I_sig = I_2
    return r_s, I_phi, I_t, I_sig


# ---------------------------------------------------------------------------
# drivers  (kerr_raytracing_ana.py:46-193)
# ---------------------------------------------------------------------------

def coords_at_tau(a, r_o, th_o, alpha, beta, tau, do_phi_and_t=True):
    """sigma and [t, r, theta, phi] at Mino times tau (kerr_raytracing_ana.py:126-193).

    tau is (G, P) or (P,); returns arrays of tau's 2-D shape.
    """
    check_params(a, r_o, th_o)
    alpha, beta = _as1d(alpha), _as1d(beta)
    if len(tau.shape) == 1:
        tau = tau.reshape(1, tau.shape[0])
    lam, eta = conserved_fixed(a, th_o, alpha, beta)
    s_o = sign_ge0(beta)
    up, um, _, _ = uplus_uminus(a, th_o, lam, eta)
    r1, r2, r3, r4, _ = radial_roots(a, lam, eta)
    th_s, G_ph, G_t = th_integrate(a, th_o, s_o, lam, eta, up, um, tau, do_phi_and_t=do_phi_and_t)
    r_s, I_ph, I_t, I_sig = r_integrate(a, r_o, lam, eta, r1, r2, r3, r4, tau, do_phi_and_t=do_phi_and_t)
    sig_s = 0 + I_sig + a ** 2 * G_t
    t_s = 0 + I_t + a ** 2 * G_t
    ph_s = 0 + I_ph + lam * G_ph
    return sig_s, [t_s, r_s, th_s, ph_s]


def mino_grid(tau_tot, ngeo, maxtaufrac=MAXTAUFRAC):
    """linspace(0, maxtaufrac*tau_tot, ngeo) (kerr_raytracing_ana.py:93-94)."""
    return np.linspace(0, maxtaufrac * tau_tot, ngeo)


def raytrace_ana(a, r_o, th_o, alpha, beta, ngeo, maxtaufrac=MAXTAUFRAC):
    """Full raytrace as kerr_raytracing_ana.py:46-122 (do_phi_and_t forced True,
    as the reference does at :97).  Returns a dict of arrays."""
    check_params(a, r_o, th_o)
    alpha, beta = _as1d(alpha), _as1d(beta)
    if len(alpha) != len(beta):
        raise Exception("alpha, beta are different lengths!")
    lam, eta = conserved_fixed(a, th_o, alpha, beta)
    r1, r2, r3, r4, rclass = radial_roots(a, lam, eta)
    tau_tot = mino_total(a, r_o, eta, r1, r2, r3, r4)
    tau = mino_grid(tau_tot, ngeo, maxtaufrac)
    sig, (t, r, th, ph) = coords_at_tau(a, r_o, th_o, alpha, beta, tau, do_phi_and_t=True)
    return dict(tau=tau, sig=sig, t=t, r=r, th=th, ph=ph, rclass=rclass,
                thclass=theta_class(eta), tau_tot=tau_tot, lam=lam, eta=eta,
                roots=(r1, r2, r3, r4))


# ---------------------------------------------------------------------------
# integer counts  (kerr_raytracing_utils.py:551-792)
# ---------------------------------------------------------------------------

def n_equatorial_crossings(a, th_o, alpha, beta, tau):
    """floor-valued count of equatorial crossings (kerr_raytracing_utils.py:605-684).
    -2 for eta<=0.  tau is (P,) or (G, P)."""
    alpha, beta = _as1d(alpha), _as1d(beta)
    one = len(tau.shape) == 1
    if one:
        tau = tau.reshape(1, tau.shape[0])
    lam, eta = conserved(a, th_o, alpha, beta)
    out = np.empty(tau.shape)
    vort = eta <= 0
    out[:, vort] = -2
    if np.any(~vort):
        bet = beta[~vort]
        tr = tau[:, ~vort]
        up, um, ratio, a2um = uplus_uminus(a, th_o, lam[~vort], eta[~vort])
        F0 = sp.ellipkinc(np.arcsin(np.cos(th_o) / np.sqrt(up)), ratio)
        K = sp.ellipkinc(0.5 * np.pi, ratio)
        w = np.sqrt(-a2um)
        res = np.empty(tr.shape)
        if th_o < np.pi / 2.:
            bm = bet <= 0
            if np.any(bm):
                res[:, bm] = (np.floor((tr * w - F0) / (2 * K)))[:, bm]
            if np.any(~bm):
                res[:, ~bm] = (np.floor((tr * w + F0) / (2 * K)) - 1)[:, ~bm]
        else:
            bm = bet >= 0
            if np.any(bm):
                res[:, bm] = (np.floor((tr * w + F0) / (2 * K)))[:, bm]
            if np.any(~bm):
                res[:, ~bm] = (np.floor((tr * w - F0) / (2 * K)) - 1)[:, ~bm]
        out[:, ~vort] = res
    return out[0] if one else out


def n_poloidal_orbits(a, th_o, alpha, beta, tau):
    """fractional poloidal orbits (kerr_raytracing_utils.py:551-603)."""
    alpha, beta = _as1d(alpha), _as1d(beta)
    one = len(tau.shape) == 1
    if one:
        tau = tau.reshape(1, tau.shape[0])
    lam, eta = conserved(a, th_o, alpha, beta)
    out = np.empty(tau.shape)
    vort = eta <= 0
    out[:, vort] = -2
    if np.any(~vort):
        up, um, ratio, a2um = uplus_uminus(a, th_o, lam[~vort], eta[~vort])
        K = sp.ellipk(ratio)
        npol = (np.sqrt(-a2um.astype(complex)) * tau[:, ~vort]) / (4 * K)
        out[:, ~vort] = np.real(npol.astype(complex))
    return out[0] if one else out


def n_angular_turnings(a, th_o, alpha, beta, tau):
    """Fractional number of angular turning points (kerr_raytracing_utils.py:686-792).

    Ordinary rays (eta > 0): (tau sqrt(-a^2 u_-) -+ F0) / (2K) + 1/2, minus F0
    for beta <= 0 (:744-749).  Rays with eta <= 0 take the reference's
    vortical formula (:753-786) as written, including its "norbit > 0.5 ->
    + 0.5" adjustment for beta <= 0 (:780-781), which the reference itself
    flags as unverified; eta == 0 rays come out NaN there (sqrt of a^2 u_- < 0).
    Like the reference, th_o must lie in [0, pi/2] (:690-691).
    """
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")
    if not (isinstance(th_o, float) and (0 <= th_o <= np.pi / 2.)):
        raise Exception("th_o should be a float in range (0,pi/2]")
    alpha, beta = _as1d(alpha), _as1d(beta)
    one = len(tau.shape) == 1
    if one:
        tau = tau.reshape(1, tau.shape[0])
    lam, eta = conserved(a, th_o, alpha, beta)
    out = np.empty(tau.shape)
    vort = eta <= 0
    out[:, vort] = 0
    with np.errstate(invalid="ignore", divide="ignore"):
        if np.any(~vort):
            bet = beta[~vort]
            tr = tau[:, ~vort]
            up, um, ratio, a2um = uplus_uminus(a, th_o, lam[~vort], eta[~vort])
            F0 = sp.ellipkinc(np.arcsin(np.cos(th_o) / np.sqrt(up)), ratio)
            K = sp.ellipkinc(0.5 * np.pi, ratio)
            w = np.sqrt(-a2um)
            res = np.empty(tr.shape)
            bm = bet <= 0
            if np.any(bm):
                res[:, bm] = ((tr * w - F0) / (2 * K))[:, bm] + 0.5
            if np.any(~bm):
                res[:, ~bm] = ((tr * w + F0) / (2 * K))[:, ~bm] + 0.5
            out[:, ~vort] = res
        if np.any(vort):
            bet = beta[vort]
            tv = tau[:, vort]
            up, um, ratio, a2um = uplus_uminus(a, th_o, lam[vort], eta[vort])
            Nu = np.arcsin(np.sqrt((np.cos(th_o) ** 2 - um) / (up - um)))
            F0 = sp.ellipkinc(Nu, 1 - ratio)
            K = sp.ellipkinc(0.5 * np.pi, 1 - ratio)
            w = np.sqrt(a2um)
            res = np.empty(tv.shape)
            bm = bet <= 0
            if np.any(bm):
                norbit = ((tv * w - F0) / (2 * K))[:, bm] + 0.5
                norbit[norbit > 0.5] += 0.5
                res[:, bm] = norbit
            if np.any(~bm):
                res[:, ~bm] = ((tv * w + F0) / (2 * K))[:, ~bm] + 0.5
            out[:, vort] = res
    return out[0] if one else out


def radial_turnings(a, th_o, alpha, beta, tau, tau_tot):
    """Radial turning points passed by Mino time tau: 1 where the ray lies
    outside the critical curve (kerr_raytracing_utils.py:794-826) and tau is
    past half its total Mino time, else 0.  A ray outside the critical curve
    reaches its one radial turning point r_4 at tau_tot / 2
    (kerr_raytracing_utils.py:518-546: tau_tot = 2 I_r(r_o)); the reference
    tests the same condition as `not (I_r <= I_max / 2)` when it picks the
    radial momentum sign (equatorial_images.py:266-288).  int8, tau's shape."""
    alpha, beta = _as1d(alpha), _as1d(beta)
    out_crit = is_outside_crit(a, th_o, alpha, beta).astype(bool)
    with np.errstate(invalid="ignore"):
        passed = ~(tau <= 0.5 * tau_tot)
    return (passed & out_crit).astype(np.int8)


# ---------------------------------------------------------------------------
# equatorial emission  (equatorial_lensing.py:85-196, equatorial_images.py,
# velocities.py:229-281, kerr_raytracing_utils.py:794-826)
# ---------------------------------------------------------------------------

def is_outside_crit(a, th_o, alpha, beta):
    """1 if outside the critical curve else 0 (kerr_raytracing_utils.py:794-826)."""
    rh = 1 + np.sqrt(1 - a ** 2)
    lam, eta = conserved(a, th_o, alpha, beta)
    vort = eta < 0
    r1, r2, r3, r4, _ = radial_roots(a, lam, eta)
    out = np.empty(alpha.shape)
    msk = (np.imag(r4) == 0) * (r4 > rh) * ~vort
    out[msk] = 1
    out[~msk] = 0
    return out


def r_equatorial(a, r_o, th_o, mbar, alpha, beta):
    """(r_s, I_r, I_max, N_max) of the mbar-th equatorial crossing
    (equatorial_lensing.py:85-196)."""
    check_params(a, r_o, th_o)
    if not (isinstance(mbar, int) and (mbar >= 0)):
        raise Exception("mbar should be an integer >=0!")
    alpha, beta = _as1d(alpha), _as1d(beta)
    lam, eta = conserved(a, th_o, alpha, beta)
    r_s = np.empty(alpha.shape)
    Ir = np.empty(alpha.shape)
    Imax = np.empty(alpha.shape)
    Nmax = np.empty(alpha.shape)
    vort = eta <= 0
    r_s[vort] = -1
    Ir[vort] = -1
    Imax[vort] = -1
    Nmax[vort] = -2
    if np.any(~vort):
        lr, er, br = lam[~vort], eta[~vort], beta[~vort]
        up, um, ratio, a2um = uplus_uminus(a, th_o, lr, er)
        F0 = sp.ellipkinc(np.arcsin(np.cos(th_o) / np.sqrt(up)), ratio)
        K = sp.ellipkinc(0.5 * np.pi, ratio)
        r1, r2, r3, r4, _ = radial_roots(a, lr, er)
        Im = mino_total(a, r_o, er, r1, r2, r3, r4)
        w = np.sqrt(-a2um)
        N = np.empty(Im.shape)
        I = np.empty(Im.shape)
        if th_o < (np.pi / 2.):
            msub = mbar + np.heaviside(br, 0)
            bm = br <= 0
            if np.any(bm):
                N[bm] = (np.floor((Im * w - F0) / (2 * K)))[bm]
                I[bm] = ((2 * msub * K + F0) / w)[bm]
            if np.any(~bm):
                N[~bm] = (np.floor((Im * w + F0) / (2 * K)) - 1)[~bm]
                I[~bm] = ((2 * msub * K - F0) / w)[~bm]
        else:
            msub = mbar + np.heaviside(-br, 0)
            bm = br >= 0
            if np.any(bm):
                N[bm] = (np.floor((Im * w + F0) / (2 * K)))[bm]
                I[bm] = ((2 * msub * K - F0) / w)[bm]
            if np.any(~bm):
                N[~bm] = (np.floor((Im * w - F0) / (2 * K)) - 1)[~bm]
                I[~bm] = ((2 * msub * K + F0) / w)[~bm]
        rs, _, _, _ = r_integrate(a, r_o, lr, er, r1, r2, r3, r4, I.reshape(1, len(I)), do_phi_and_t=False)
        rs = rs[0]
        rs[N < mbar] = 0
        r_s[~vort] = rs
        Ir[~vort] = I
        Imax[~vort] = Im
        Nmax[~vort] = N
    return r_s, Ir, Imax, Nmax


# ---------------------------------------------------------------------------
# lensed curves of constant equatorial radius  (equatorial_lensing.py:198-293)
# ---------------------------------------------------------------------------

LENS_INF = 1.e50      # equatorial_lensing.py:18
LENS_R0 = np.inf      # equatorial_lensing.py:19


def r_equatorial2(rho, varphi, a, th0, mbar=0):
    """r_s of subring mbar at the screen point rho (cos varphi, sin varphi),
    1e50 for rays outside the critical curve with no such crossing, 1 for
    rays inside it (equatorial_lensing.py:198-216)."""
    alpha = np.cos(varphi) * rho
    beta = np.sin(varphi) * rho
    with np.errstate(all="ignore"):
        r, Ir, Imax, Nmax = r_equatorial(a, LENS_R0, th0, mbar, np.array([alpha]), np.array([beta]))
    r = r[0]
    outside = is_outside_crit(a, th0, np.array([alpha]), np.array([beta]))[0]
    if r <= 0 and outside == 1:
        r = LENS_INF
    elif r <= 0:
        r = 1
    return r


def rho_bracket(req, mbar):
    """The bracket rho_of_req searches (equatorial_lensing.py:229-238)."""
    if mbar == 0:
        return 0, req + 5
    if mbar == 1:
        return 0, 100
    return 1, 10


def rho_of_req_single(a, th0, req, varphi, mbar=0):
    """brentq of r_equatorial2 - req over the bracket (equatorial_lensing.py:249-262)."""
    from scipy.optimize import brentq
    lo, hi = rho_bracket(req, mbar)
    return brentq(lambda rho: r_equatorial2(rho, varphi, a, th0, mbar) - req, lo, hi)


def rho_of_req(a, th0, req, mbar=0, varphis=None):
    """(varphis, rhos, alphas, betas) over 360 angles by default
    (equatorial_lensing.py:226-247)."""
    if varphis is None:
        varphis = np.linspace(-180, 179, 360) * np.pi / 180.
    rhos = np.array([rho_of_req_single(a, th0, req, vp, mbar) for vp in varphis])
    return varphis, rhos, np.cos(varphis) * rhos, np.sin(varphis) * rhos


def critical_curve(a, th0, n=100000):
    """Screen points of the bound photon orbits r in [r_-, r_+], NaNs
    dropped, mirrored in beta (equatorial_lensing.py:264-293)."""
    rminus = 2 * (1 + np.cos(2. * np.arccos(-a) / 3.))
    rplus = 2 * (1 + np.cos(2. * np.arccos(a) / 3.))
    r = np.linspace(rminus, rplus, n)
    with np.errstate(invalid="ignore"):
        lam = -(r ** 3 + a ** 2 + r * (a ** 2) - 3 * r ** 2) / (a * r - a)
        eta = (r ** 2) * (3 * r ** 2 + a ** 2 - lam ** 2) / (r ** 2 - a ** 2)
        alpha = -lam / np.sin(th0)
        beta = np.sqrt(eta + (a * np.cos(th0)) ** 2 - (lam / np.tan(th0)) ** 2)
    mask = np.isnan(alpha) + np.isnan(beta)
    alpha, beta = alpha[~mask], beta[~mask]
    return np.hstack((alpha, np.flip(alpha))), np.hstack((beta, -np.flip(beta)))


def r_isco(a, retrograde=False):
    """ISCO radius (velocities.py:242-244)."""
    s = -1 if retrograde else 1
    z1 = 1 + np.cbrt(1 - a ** 2) * (np.cbrt(1 + a) + np.cbrt(1 - a))
    z2 = np.sqrt(3 * a ** 2 + z1 ** 2)
    return 3 + z2 - s * np.sqrt((3 - z1) * (3 + z1 + 2 * z2))


def u_kep(a, r, retrograde=False):
    """Keplerian outside the ISCO, Cunningham plunge inside (velocities.py:229-281).
    Returns (u0, u1, u2, u3) with u2 = 0."""
    r = _as1d(r)
    s = -1 if retrograde else 1
    ri = r_isco(a, retrograde)
    u0 = np.zeros(r.shape)
    u1 = np.zeros(r.shape)
    u3 = np.zeros(r.shape)
    out = r >= ri
    spin = np.abs(a)
    asign = np.sign(a)
    if np.any(out):
        rr = r[out]
        Om = asign * s / (rr ** 1.5 + s * spin)
        u0[out] = (rr ** 1.5 + s * spin) / np.sqrt(rr ** 3 - 3 * rr ** 2 + 2 * s * spin * rr ** 1.5)
        u3[out] = Om * u0[out]
    if np.any(~out):
        rr = r[~out]
        Delta = (rr ** 2 - 2 * rr + a ** 2)
        ell = s * asign * (ri ** 2 + a ** 2 - s * 2 * spin * np.sqrt(ri)) / (ri ** 1.5 - 2 * np.sqrt(ri) + s * spin)
        gam = np.sqrt(1 - 2. / (3. * ri))
        H = (2 * rr - a * ell) / Delta
        u0[~out] = gam * (1 + (2 / rr) * (1 + H))
        u1[~out] = -np.sqrt(2. / (3 * ri)) * (ri / rr - 1) ** 1.5
        u3[~out] = gam * (ell + a * H) / (rr ** 2)
    return u0, u1, 0, u3


def radial_momentum_sign(a, th_o, alpha, beta, Ir, Imax):
    """+1 outgoing (inside the critical curve or before the turning point),
    -1 otherwise (equatorial_images.py:266-288)."""
    inside = ~is_outside_crit(a, th_o, alpha, beta).astype(bool)
    sign = np.empty(alpha.shape)
    outgoing = inside + (Ir <= 0.5 * Imax)
    sign[outgoing] = 1
    sign[~outgoing] = -1
    return sign


def theta_momentum_sign(th_o, mbar):
    """(equatorial_images.py:290-305)."""
    if th_o < 0.5 * np.pi:
        return -1 * np.power(-1, np.mod(mbar, 2))
    return 1 * np.power(-1, np.mod(mbar, 2))


def calc_redshift(a, r, lam, eta, kr_sign, kth_sign, u0, u1, u2, u3, th=np.pi / 2):
    """g = 1/(-k.u) with k from (lam, eta) (equatorial_images.py:307-334)."""
    a2 = a ** 2
    r2 = r ** 2
    cth2 = np.cos(th) ** 2
    sth2 = np.sin(th) ** 2
    Delta = r2 - 2 * r + a2
    R = (r2 + a2 - a * lam) ** 2 - Delta * (eta + (lam - a) ** 2)
    TH = eta + a2 * cth2 - lam * lam * cth2 / sth2
    return 1 / (1 * u0 - lam * u3 - np.sign(kth_sign) * u2 * np.sqrt(TH) - np.sign(kr_sign) * u1 * np.sqrt(R) / Delta)


def equatorial_subring(a, r_o, th_o, mbar, alpha, beta, p=2.0):
    """One subring of the equatorial image: g^3 r_s^-p on unmasked pixels.

    Composition of equatorial_images.py:160-196 (masks :175, momentum signs
    :188-189, u_kep velocity :194, redshift :195) with the emissivity r_s^-p
    and the g^3 weighting of the workload definition.  Returns
    (I, g, r_s, Ir, Imax, Nmax, zeromask).
    """
    alpha, beta = _as1d(alpha), _as1d(beta)
    rh = 1 + np.sqrt(1 - a ** 2)
    lam, eta = conserved(a, th_o, alpha, beta)
    r_s, Ir, Imax, Nmax = r_equatorial(a, r_o, th_o, mbar, alpha, beta)
    g = np.zeros(alpha.shape)
    I = np.zeros(alpha.shape)
    zm = (Nmax < mbar) + (Nmax == -1) + (r_s <= rh)
    if np.any(~zm):
        krs = radial_momentum_sign(a, th_o, alpha[~zm], beta[~zm], Ir[~zm], Imax[~zm])
        kths = theta_momentum_sign(th_o, mbar)
        u0, u1, u2, u3 = u_kep(a, r_s[~zm])
        gg = calc_redshift(a, r_s[~zm], lam[~zm], eta[~zm], krs, kths, u0, u1, u2, u3, th=np.pi / 2)
        g[~zm] = gg
        I[~zm] = gg ** 3 * r_s[~zm] ** (-p)
    return I, g, r_s, Ir, Imax, Nmax, zm


def equatorial_image(a, r_o, th_o, alpha, beta, mbar_max=2, p=2.0):
    """Sum over subrings mbar = 0..mbar_max of g^3 r_s^-p.  Returns
    (image, per-subring dict)."""
    img = np.zeros(_as1d(alpha).shape)
    parts = []
    for mbar in range(mbar_max + 1):
        I, g, r_s, Ir, Imax, Nmax, zm = equatorial_subring(a, r_o, th_o, mbar, alpha, beta, p)
        img = img + I
        parts.append(dict(I=I, g=g, r_s=r_s, Ir=Ir, Imax=Imax, Nmax=Nmax, zeromask=zm))
    return img, parts
