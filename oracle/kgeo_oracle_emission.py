"""CPU oracle for the equatorial emission models and the polarized image
(TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline may import this module; the product never does).

A NumPy/SciPy/mpmath restatement of the reference's emission pipeline:

  velocities   kgeo/velocities.py:23-698   (Velocity, u_lab, tetrades, models)
  B fields     kgeo/bfields.py:7-806       (Bfield, bfield_lab, omega_field,
                                             bfield_fluid, bsq, field models)
  emissivity   kgeo/emissivities.py:37-157 (Emissivity, jrest, models)
  image        kgeo/equatorial_images.py:23-544 (make_image, Iobs,
                                             calc_polquantities, calc_evpa,
                                             calc_pathlength_equatorial)

Deviations from the reference, each deliberate:
  * bz_monopole's f(r) is the reference's defining formula (bfields.py:7-13,
    mpmath polylog) and f'(r) its derivative (mpmath.diff); the reference
    itself interpolates a table of that formula (bz_fr_data.dat, :23-30),
    which is not part of the reference tree.
  * driftframe supports numeric nu_parallel without gammamax; 'FF' needs the
    force-free stagnation-surface solver (ff_boost.py), which is out of scope.
  * make_image(nmax_only=True) returns its arrays (the reference's returns an
    unassigned name).

How each part is pinned is listed in tests/test_oracle_emission.py.
"""

import mpmath as mp
import numpy as np
import scipy.special as sp

from .kgeo_oracle import (_as1d, conserved, is_outside_crit, n_poloidal_orbits, r_equatorial, radial_momentum_sign,
                          theta_momentum_sign, u_kep)

# ---------------------------------------------------------------------------
# metric (Boyer-Lindquist), as the reference writes it in every model
# ---------------------------------------------------------------------------


def metric(a, r, th=np.pi / 2):
    """Covariant and contravariant BL metric components
    (equatorial_images.py:351-369)."""
    a2 = a ** 2
    r2 = r ** 2
    cth2 = np.cos(th) ** 2
    sth2 = np.sin(th) ** 2
    Delta = r2 - 2 * r + a2
    Sigma = r2 + a2 * cth2
    m = dict(a2=a2, r2=r2, cth2=cth2, sth2=sth2, Delta=Delta, Sigma=Sigma)
    m["g00"] = -(1 - 2 * r / Sigma)
    m["g11"] = Sigma / Delta
    m["g22"] = Sigma
    m["g33"] = (r2 + a2 + 2 * r * a2 * sth2 / Sigma) * sth2
    m["g03"] = -2 * r * a * sth2 / Sigma
    m["gu00"] = -(r2 + a2 + 2 * r * a2 * sth2 / Sigma) / Delta
    m["gu11"] = Delta / Sigma
    m["gu22"] = 1. / Sigma
    m["gu33"] = (Delta - a2 * sth2) / (Sigma * Delta * sth2)
    m["gu03"] = -(2 * a * r) / (Sigma * Delta)
    return m


def lower(m, u):
    u0, u1, u2, u3 = u
    return (m["g00"] * u0 + m["g03"] * u3, m["g11"] * u1, m["g22"] * u2, m["g33"] * u3 + m["g03"] * u0)


def _check_spin(a):
    if not (isinstance(a, float) and (0 <= np.abs(a) < 1)):
        raise Exception("|a| should be a float in range [0,1)")


def _isco(a, s):
    z1 = 1 + np.cbrt(1 - a ** 2) * (np.cbrt(1 + a) + np.cbrt(1 - a))
    z2 = np.sqrt(3 * a ** 2 + z1 ** 2)
    return 3 + z2 - s * np.sqrt((3 - z1) * (3 + z1 + 2 * z2))


# ---------------------------------------------------------------------------
# B fields (bfields.py)
# ---------------------------------------------------------------------------

def bz_f_exact(r):
    """bfields.py:7-13: the BZ monopole radial function, real part, at 40
    digits (its terms cancel to f ~ 1/(4r): about r^3 / f of the working
    precision is lost); at r = 2 the limit, ln(1 - 2/r) ln(r/2) -> 0."""
    with mp.workdps(40):
        return float(mp.re(_bz_f_mp(mp.mpf(r))))


RMAXINTERP = 100.   # bfields.py:24-26: beyond it f = 1/(4r)


def bz_f(r):
    r = _as1d(r)
    out = np.array([bz_f_exact(x) if x <= RMAXINTERP else 1. / (4. * x) for x in r])
    return out


def bz_fprime(r):
    r = _as1d(r)
    with mp.workdps(40):
        return np.array([float(mp.diff(lambda x: mp.re(_bz_f_mp(x)), mp.mpf(x))) if x <= RMAXINTERP
                         else -1. / (4. * x * x) for x in r])


def _bz_f_mp(r):
    if r == 2:
        A = mp.pi ** 2 / 6
    else:
        rc = r + 0j
        A = mp.polylog(2, 2 / rc) - mp.log(1 - 2 / rc) * mp.log(rc / 2)
    r = r + 0j
    out = A * r * r * (2 * r - 3) / 8
    out += (1 + 3 * r - 6 * r * r) * mp.log(r / 2) / 12
    return out + mp.mpf(11) / 72 + 1 / (3 * r) + r / 2 - r * r / 2


def Bfield_simple(a, r, coeffs, rm1=False):
    """bfields.py:429-477 (rm1: Bfield_simple_rm1, Br ~ r/gdet)."""
    _check_spin(a)
    arad, avert, ator = coeffs
    th = np.pi / 2.
    Sigma = r ** 2 + a ** 2 * np.cos(th) ** 2
    gdet = np.abs(np.sin(th)) * Sigma
    rad = (r * np.sin(th) / gdet) if rm1 else (np.sin(th) / gdet)
    Br = arad * rad + avert * (r * np.cos(th) / gdet)
    Bth = avert * (-np.sin(th) / gdet)
    Bph = ator * (1. / gdet)
    return Br, Bth, Bph


def Bfield_monopoleA(a, r, th, C=1, omegafac=0.5):
    """bfields.py:480-501."""
    rh = 1 + np.sqrt(1 - a ** 2)
    omegah = a / (2 * rh)
    phih = 4 * np.pi * C
    Delta = r ** 2 - 2 * r + a ** 2
    Sigma = r ** 2 + a ** 2 * np.cos(th) ** 2
    omegaf = omegafac * omegah
    I = phih * (np.sin(th) ** 2) * rh * (omegaf - omegah) / (rh ** 2 + a ** 2 * np.cos(th) ** 2)
    B1 = (phih / (4 * np.pi)) / Sigma
    B3 = I / (2 * np.pi * Delta * (np.sin(th) ** 2))
    return B1, 0 * B1, B3, omegaf


def Bfield_BZmagic(a, r, th, C=1):
    """bfields.py:504-542 ('bz_guess')."""
    _check_spin(a)
    a2 = a ** 2
    r2 = r ** 2
    sth2 = np.sin(th) ** 2
    Delta = r2 - 2 * r + a2
    Pi = (r2 + a2) ** 2 - a2 * Delta * sth2
    rH = 1 + np.sqrt(1 - a2)
    OmegaH = a / (2 * rH)
    spin = np.abs(a)
    f2 = (6 * np.pi * np.pi - 49.) / 72.
    omega2 = 1. / 32. - (4 * f2 - 1) * sth2 / 64.
    OmegaBZ = spin * (1. / 8.) + (spin ** 3) * omega2
    if a < 0:
        OmegaBZ = -OmegaBZ
    Br = C / r2
    Bph = -1 * (OmegaH - OmegaBZ) * np.sqrt(Pi) / Delta * Br
    return Br, 0 * Br, Bph, OmegaBZ


def Bfield_BZmonopole(a, r, th, C=1, secondorder_only=False):
    """bfields.py:545-592, with f(r) and f'(r) evaluated exactly."""
    _check_spin(a)
    r = _as1d(r)
    a2 = a ** 2
    r2 = r ** 2
    sth = np.sin(th)
    cth = np.cos(th)
    cth2 = cth ** 2
    sth2 = sth ** 2
    Delta = r2 - 2 * r + a2
    Sigma = r2 + a2 * cth2
    gdet = sth * Sigma
    fr = bz_f(r)
    frp = bz_fprime(r)
    dphidtheta = C * sth * (1 + 0.5 * a2 * fr * (1 + 3 * np.cos(2 * th)))
    dphidr = C * a2 * sth2 * cth * frp
    f2 = (6 * np.pi * np.pi - 49.) / 72.
    omega2 = 1. / 32. - sth2 * (4 * f2 - 1) / 64.
    I = -1 * C * 2 * np.pi * sth2 * (a / 8. + (a ** 3) * (omega2 + 0.25 * fr * cth2))
    Br = dphidtheta / gdet
    Bth = -dphidr / gdet
    Bph = I / (2 * np.pi * Delta * sth2)
    if secondorder_only:
        Br = C / r2 + C * a2 * (-cth2 / (r2 * r2) + fr * (0.5 + 1.5 * np.cos(2 * th)) / r2)
        Bth = -C * a2 * frp * cth * sth / r2
        Bph = -C * 0.125 * a / (r2 - 2 * r) + C * (a ** 3) * (0.125 / ((r2 - 2 * r) ** 2)
                                                              - (omega2 + 0.25 * cth2 * fr) / (r2 - 2 * r))
    spin = np.abs(a)
    OmegaBZ = spin / 8. + spin ** 3 * omega2
    if a < 0:
        OmegaBZ = -OmegaBZ
    return Br, Bth, Bph, OmegaBZ


def omega_BZpara(th, psi, a):
    """bfields.py:609-615 (real branch of Lambert W)."""
    rp = 1 + np.sqrt(1 - a ** 2)
    w = sp.lambertw(-psi / rp + np.log(4), k=0)
    x = np.exp(w)
    y = 1 + psi / (1 + w) / (x * (2 - x))
    return a / 4 / y


def Bfield_BZpara(a, r, th, C=1):
    """bfields.py:627-673."""
    _check_spin(a)
    r = _as1d(r)
    a2 = a ** 2
    r2 = r ** 2
    rp = 1 + np.sqrt(1 - a2)
    sth = np.sin(th)
    cth = np.cos(th)
    abscth = np.abs(cth)
    Delta = r2 - 2 * r + a2
    Sigma = r2 + a2 * cth ** 2
    gdet = sth * Sigma
    psi = r * (1 - abscth) + rp * (1 + abscth) * (1 - np.log(1 + abscth)) - 2 * rp * (1 - np.log(2))
    dpsidtheta = np.sign(cth) * sth * (r + rp * np.log(1 + abscth))
    dpsidr = 1 - abscth
    argvals = -psi / rp + np.log(4)
    ok = argvals >= -1 / np.e
    Om = np.zeros(r.shape)
    Om[ok] = np.real(omega_BZpara(th, psi[ok], a))
    OmegaBZ = C * Om
    I = -4 * np.pi * psi * OmegaBZ * np.sign(cth)
    Br = C * dpsidtheta / gdet
    Bth = -C * dpsidr / gdet
    Bph = I / (2 * np.pi * Delta * sth ** 2)
    return Br, Bth, Bph, OmegaBZ


def omega_BZpower(th, psi, a, p, usemono=False):
    """bfields.py:618-624."""
    if usemono or p == 0:
        return a / 8
    rp = 1 + np.sqrt(1 - a ** 2)
    cthhorizon = 1 - psi * rp ** (-p)
    return a / (4 + 8 / (1 + cthhorizon))


def Bfield_power(a, r, th, p, C=1, usemono=False):
    """bfields.py:676-709."""
    _check_spin(a)
    a2 = a ** 2
    r2 = r ** 2
    sth = np.sin(th)
    cth = np.cos(th)
    abscth = np.abs(cth)
    Delta = r2 - 2 * r + a2
    Sigma = r2 + a2 * cth ** 2
    gdet = sth * Sigma
    psi = r ** p * (1 - abscth)
    dpsidtheta = np.sign(cth) * sth * (r ** p)
    dpsidr = p * psi / r
    Om = omega_BZpower(th, psi, a, p, usemono=usemono)
    if p > 0:
        I = -4 * np.pi * psi * Om * np.sign(cth)
    else:
        I = -2 * np.pi * psi * (2 - psi) * Om * np.sign(cth)
    Br = C * dpsidtheta / gdet
    Bth = -C * dpsidr / gdet
    Bph = I / (2 * np.pi * Delta * sth ** 2)
    return Br, Bth, Bph, Om


def Bfield_gen_power(a, r, th, n_I, p_val, isAbove):
    """bfields.py:712-746."""
    a2 = a ** 2
    r2 = r ** 2
    rp = 1 + np.sqrt(1 - a2)
    sth = np.sin(th)
    cth = np.cos(th)
    Delta = r2 - 2 * r + a2
    Sigma = r2 + a2 * cth ** 2
    gdet = sth * Sigma
    Omega_H = a / (2 * rp)
    sgn = 1 if isAbove else -1
    I_0 = sgn * (a * np.pi * sth ** 2) / (rp * rp + a2 * cth ** 2)
    dpsidtheta = sgn * sth * ((r / rp) ** p_val)
    I = I_0 * (r / rp) ** (-1 * n_I)
    dpsidr = (1 - np.abs(cth)) * ((p_val * (r ** (p_val - 1))) / (rp ** p_val))
    Br = dpsidtheta / gdet
    Bth = -1 * (dpsidr / gdet)
    Bph = I / (2 * np.pi * Delta * sth ** 2)
    return Br, Bth, Bph, 0.5 * Omega_H


def load_table(path, names):
    """The reference's '# r, ...' CSV tables (velocities.py:626-655,
    bfields.py:749-787): columns r and names (zeros where absent)."""
    with open(path) as f:
        header = f.readline().strip()
    if header[0] != "#":
        raise Exception("file %s does not have a header!" % path)
    header = [x.strip() for x in header[1:].split(",")]
    data = np.loadtxt(path, delimiter=",", skiprows=1)
    radii = data[:, header.index("r")]
    cols = [data[:, header.index(n)] if n in header else np.zeros_like(radii) for n in names]
    return radii, cols


class Bfield:
    """bfields.py:33-91 (constructor), :93-131 (bfield_lab), :175-196
    (omega_field), :330-395 (bfield_fluid, bsq)."""

    def __init__(self, fieldtype="rad", **kw):
        self.fieldtype = fieldtype
        lab = ["rad", "vert", "tor", "simple", "simple_rm1", "monopoleA", "bz_monopole", "bz_guess", "bz_para",
               "power", "fromfile", "gen_power"]
        if fieldtype in lab:
            self.fieldframe = "lab"
        elif fieldtype == "const_comoving":
            self.fieldframe = "comoving"
        else:
            raise Exception("fieldtype %s not recognized in Bfield!" % fieldtype)
        self.kw = kw
        if fieldtype in ["bz_monopole", "bz_guess", "bz_para"]:
            self.secondorder_only = kw.get("secondorder_only", False)
            self.C = kw.get("C", 1)
        elif fieldtype == "monopoleA":
            self.omegafac = kw.get("omegafac", 0.5)
            self.C = kw.get("C", 1)
        elif fieldtype == "power":
            self.secondorder_only = kw.get("secondorder_only", False)
            self.C = kw.get("C", 1)
            self.pval = kw.get("p", 1)
            self.usemono = kw.get("usemono", False)
        elif fieldtype in ("rad", "vert", "tor"):
            self.Cr, self.Cvert, self.Cph = {"rad": (1, 0, 0), "vert": (0, 1, 0), "tor": (0, 0, 1)}[fieldtype]
        elif fieldtype == "fromfile":
            self.radii, (self.tBr, self.tBth, self.tBph) = load_table(kw.get("file"), ["Br", "Btheta", "Bphi"])
        elif fieldtype == "gen_power":
            self.n_I = kw.get("n_I", 0)
            self.p_val = kw.get("p_val", 1)
            self.isAbove = kw.get("isAbove", True)
        else:
            self.Cr = kw.get("Cr", 0)
            self.Cvert = kw.get("Cvert", 0)
            self.Cph = kw.get("Cph", 0)
            if self.Cr == self.Cvert == self.Cph == 0.:
                raise Exception("all field coefficients are 0!")

    def _full(self, a, r, th):
        t = self.fieldtype
        if t == "bz_monopole":
            return Bfield_BZmonopole(a, r, th, self.C, secondorder_only=self.secondorder_only)
        if t == "monopoleA":
            return Bfield_monopoleA(a, r, th, self.C, omegafac=self.omegafac)
        if t == "bz_para":
            return Bfield_BZpara(a, r, th, self.C)
        if t == "bz_guess":
            return Bfield_BZmagic(a, r, th, self.C)
        if t == "power":
            return Bfield_power(a, r, th, self.pval, C=self.C, usemono=self.usemono)
        if t == "gen_power":
            return Bfield_gen_power(a, r, th, self.n_I, self.p_val, self.isAbove)
        return None

    def bfield_lab(self, a, r, th=np.pi / 2):
        r = _as1d(r)
        if self.fieldframe != "lab":
            raise Exception("Bfield.bfield_lab only supported for Bfield.fieldtype==lab")
        t = self.fieldtype
        if t in ["simple", "rad", "vert", "tor"]:
            return Bfield_simple(a, r, (self.Cr, self.Cvert, self.Cph))
        if t == "simple_rm1":
            return Bfield_simple(a, r, (self.Cr, self.Cvert, self.Cph), rm1=True)
        if t == "fromfile":
            return tuple(np.interp(r, self.radii, c) for c in (self.tBr, self.tBth, self.tBph))
        return self._full(a, r, th)[:3]

    def bfield_comoving(self, a, r):
        r = _as1d(r)
        if self.fieldframe != "comoving":
            raise Exception("Bfield.bfield_comoving only supported for Bfield.fieldtype==comoving")
        return (self.Cr * np.ones(r.shape), -1 * self.Cvert * np.ones(r.shape), self.Cph * np.ones(r.shape))

    def omega_field(self, a, r, th=np.pi / 2):
        r = _as1d(r)
        full = self._full(a, r, th)
        if full is None:
            raise Exception("self.omega_field not implemented for fieldtype %s'!" % self.fieldtype)
        return full[3]

    def bfield_fluid(self, a, r, velocity, th=np.pi / 2.):
        r = _as1d(r)
        m = metric(a, r, th)
        B1, B2, B3 = self.bfield_lab(a, r, th=th)
        u = velocity.u_lab(a, r, th=th)
        u0, u1, u2, u3 = u
        _, u1_l, u2_l, u3_l = lower(m, u)
        b0 = B1 * u1_l + B2 * u2_l + B3 * u3_l
        return (b0, (B1 + b0 * u1) / u0, (B2 + b0 * u2) / u0, (B3 + b0 * u3) / u0)

    def bsq(self, a, r, velocity, th=np.pi / 2.):
        r = _as1d(r)
        m = metric(a, r, th)
        b = self.bfield_fluid(a, r, velocity, th=th)
        bl = lower(m, b)
        return b[0] * bl[0] + b[1] * bl[1] + b[2] * bl[2] + b[3] * bl[3]


# ---------------------------------------------------------------------------
# velocities (velocities.py)
# ---------------------------------------------------------------------------

def u_zamo(a, r):
    """velocities.py:179-208."""
    _check_spin(a)
    r = _as1d(r)
    th = np.pi / 2.
    Sigma = r ** 2 + a ** 2 * np.cos(th) ** 2
    g00 = -(1 - 2 * r / Sigma)
    g33 = (r ** 2 + a ** 2 + 2 * r * (a * np.sin(th)) ** 2 / Sigma) * np.sin(th) ** 2
    g03 = -2 * r * a * np.sin(th) ** 2 / Sigma
    v3 = -g03 / g33
    u0 = np.sqrt(-1. / (g00 + 2 * g03 * v3 + g33 * v3 * v3))
    return u0, 0, 0, u0 * v3


def u_infall(a, r):
    """velocities.py:211-226."""
    _check_spin(a)
    r = _as1d(r)
    Delta = r ** 2 + a ** 2 - 2 * r
    Xi = (r ** 2 + a ** 2) ** 2 - Delta * a ** 2
    return Xi / (r ** 2 * Delta), -np.sqrt(2 * r * (r ** 2 + a ** 2)) / (r ** 2), 0, 2 * a / (r * Delta)


def u_subkep(a, r, fac_subkep=1, retrograde=False):
    """velocities.py:284-359."""
    _check_spin(a)
    if not (0 <= fac_subkep <= 1):
        raise Exception("fac_subkep should be in the range [0,1]")
    r = _as1d(r)
    s = -1 if retrograde else 1
    ri = _isco(a, s)
    u0, u1, u3 = np.zeros(r.shape), np.zeros(r.shape), np.zeros(r.shape)
    out = r >= ri
    spin = np.abs(a)
    asign = np.sign(a)

    def ell_gam(rr):
        Delta = rr ** 2 - 2 * rr + a ** 2
        Xi = (rr ** 2 + a ** 2) ** 2 - Delta * a ** 2
        ell = asign * s * (rr ** 2 + spin ** 2 - s * 2 * spin * np.sqrt(rr)) / (rr ** 1.5 - 2 * np.sqrt(rr) + s * spin)
        ell = ell * fac_subkep
        return ell, np.sqrt(Delta / (Xi / rr ** 2 - 4 * a * ell / rr - (1 - 2 / rr) * ell ** 2))

    if np.any(out):
        rr = r[out]
        ell, gam = ell_gam(rr)
        Delta = rr ** 2 - 2 * rr + a ** 2
        H = (2 * rr - a * ell) / Delta
        chi = 1 / (1 + (2 / rr) * (1 + H))
        Om = (chi / rr ** 2) * (ell + a * H)
        u0[out] = gam / chi
        u3[out] = (gam / chi) * Om
    if np.any(~out):
        rr = r[~out]
        Delta = rr ** 2 - 2 * rr + a ** 2
        Xi = (rr ** 2 + a ** 2) ** 2 - Delta * a ** 2
        ell_i, gam_i = ell_gam(ri)
        H = (2 * rr - a * ell_i) / Delta
        chi = 1 / (1 + (2 / rr) * (1 + H))
        Om = (chi / rr ** 2) * (ell_i + a * H)
        nu = (rr / Delta) * np.sqrt(Xi / rr ** 2 - 4 * a * ell_i / rr - (1 - 2 / rr) * ell_i ** 2 - Delta / gam_i ** 2)
        u0[~out] = gam_i / chi
        u1[~out] = -gam_i * (Delta / rr ** 2) * nu
        u3[~out] = (gam_i / chi) * Om
    return u0, u1, 0, u3


def u_gelles(a, r, beta=0.3, chi=-150 * (np.pi / 180.)):
    """velocities.py:362-389."""
    _check_spin(a)
    r = _as1d(r)
    a2 = a ** 2
    r2 = r ** 2
    th = np.pi / 2.
    sth2 = np.sin(th) ** 2
    Delta = r2 - 2 * r + a2
    Xi = (r2 + a2) ** 2 - Delta * a2 * sth2
    omegaz = 2 * a * r / Xi
    gamma = 1 / np.sqrt(1 - beta ** 2)
    u0 = (gamma / r) * np.sqrt(Xi / Delta)
    u1 = (beta * gamma * np.cos(chi) / r) * np.sqrt(Delta)
    u3 = (gamma * omegaz / r) * np.sqrt(Xi / Delta) + (r * beta * gamma * np.sin(chi)) / np.sqrt(Xi)
    return u0, u1, 0, u3


def u_grmhd_fit(a, r, ell_isco=1., vr_isco=2, p1=6., p2=2., dd=0.2):
    """velocities.py:392-439 ('simfit')."""
    _check_spin(a)
    r = _as1d(r)
    r_isco = _isco(a, 1)
    m = metric(a, r, np.pi / 2.)
    ell = ell_isco * (r / r_isco) ** .5
    vr = -vr_isco * ((r / r_isco) ** (-p1)) * (0.5 * (1 + (r / r_isco) ** (1 / dd))) ** ((p1 - p2) * dd)
    gam = np.sqrt(-1. / (m["gu00"] + m["gu11"] * vr * vr + m["gu33"] * ell * ell - 2 * m["gu03"] * ell))
    u_0, u_1, u_3 = -gam, vr * gam, ell * gam
    return (m["gu00"] * u_0 + m["gu03"] * u_3, m["gu11"] * u_1, 0, m["gu33"] * u_3 + m["gu03"] * u_0)


def u_general(a, r, fac_subkep=1, beta_phi=1, beta_r=1, retrograde=False):
    """velocities.py:442-471."""
    _check_spin(a)
    if not (0 <= fac_subkep <= 1):
        raise Exception("fac_subkep should be in the range [0,1]")
    if not (0 <= beta_phi <= 1):
        raise Exception("beta_phi should be in the range [0,1]")
    if not (0 <= beta_r <= 1):
        raise Exception("beta_r should be in the range [0,1]")
    r = _as1d(r)
    Delta = r ** 2 + a ** 2 - 2 * r
    u0i, u1i, _, u3i = u_infall(a, r)
    u0s, u1s, _, u3s = u_subkep(a, r, retrograde=retrograde, fac_subkep=fac_subkep)
    u1 = u1s + (1 - beta_r) * (u1i - u1s)
    Om = u3s / u0s + (1 - beta_phi) * (u3i / u0i - u3s / u0s)
    u0 = np.sqrt(1 + (r ** 2) * (u1 ** 2) / Delta)
    u0 = u0 / np.sqrt(1 - (r ** 2 + a ** 2) * Om ** 2 - (2 / r) * (1 - a * Om) ** 2)
    return u0, u1, 0, u0 * Om


def u_driftframe(a, r, bfield, nu_parallel=0, th=np.pi / 2):
    """velocities.py:487-623 with numeric nu_parallel and no gammamax."""
    nu_parallel = nu_parallel * np.ones_like(r)
    _check_spin(a)
    if np.any(np.logical_or(nu_parallel > 1, nu_parallel < -1)):
        raise Exception("nu_parallel should be in the range (-1,1)")
    r = _as1d(r)
    a2 = a ** 2
    r2 = r ** 2
    cth2 = np.cos(th) ** 2
    sth2 = np.sin(th) ** 2
    Delta = r2 - 2 * r + a2
    Sigma = r2 + a2 * cth2
    Xi = (r2 + a2) ** 2 - Delta * a2 * sth2
    omegaz = 2 * a * r / Xi
    gdet = Sigma * np.sin(th)
    g11, g22, g33 = Sigma / Delta, Sigma, Xi * sth2 / Sigma
    alpha = np.sqrt(Delta * Sigma / Xi)
    eta3 = 2 * a * r / np.sqrt(Delta * Sigma * Xi)
    omegaf = bfield.omega_field(a, r, th=th)
    B1, B2, B3 = bfield.bfield_lab(a, r, th=th)
    E1 = (omegaf - omegaz) * Xi * np.sin(th) * B2 / Sigma
    E2 = -(omegaf - omegaz) * Xi * np.sin(th) * B1 / (Sigma * Delta)
    E3 = 0
    Bsq = g11 * B1 * B1 + g22 * B2 * B2 + g33 * B3 * B3
    Esq = g11 * E1 * E1 + g22 * E2 * E2 + g33 * E3 * E3
    B1c, B2c, B3c = g11 * B1, g22 * B2, g33 * B3
    E1c, E2c, E3c = g11 * E1, g22 * E2, g33 * E3
    k = alpha / (Bsq * gdet)
    vperp1 = k * (E2c * B3c - B2c * E3c)
    vperp2 = k * (E3c * B1c - B3c * E1c)
    vperp3 = k * (E1c * B2c - B1c * E2c)
    vpar_max = np.sqrt(1 - Esq / Bsq)
    sB = np.sqrt(Bsq)
    v1 = vperp1 + nu_parallel * vpar_max * B1 / sB
    v2 = vperp2 + nu_parallel * vpar_max * B2 / sB
    v3 = vperp3 + nu_parallel * vpar_max * B3 / sB
    vsq = g11 * v1 * v1 + g22 * v2 * v2 + g33 * v3 * v3
    gamma = 1. / np.sqrt(1 - vsq)
    return gamma / alpha, gamma * v1, gamma * v2, gamma * (v3 + eta3)


class Velocity:
    """velocities.py:23-176."""

    def __init__(self, veltype="kep", **kw):
        self.veltype = veltype
        self.kw = kw
        if veltype in ("zamo", "infall"):
            pass
        elif veltype in ("kep", "cunningham"):
            self.retrograde = kw.get("retrograde", False)
        elif veltype in ("subkep", "cunningham_subkep"):
            self.retrograde = kw.get("retrograde", False)
            self.fac_subkep = kw.get("fac_subkep", 1)
        elif veltype == "general":
            self.retrograde = kw.get("retrograde", False)
            self.fac_subkep = kw.get("fac_subkep", 1)
            self.beta_phi = kw.get("beta_phi", 1)
            self.beta_r = kw.get("beta_r", 1)
        elif veltype == "gelles":
            self.gelles_beta = kw.get("gelles_beta", 0.3)
            self.gelles_chi = kw.get("gelles_chi", -150 * np.pi / 180.)
        elif veltype == "simfit":
            self.ell_isco = kw.get("ell_isco", 1.)
            self.vr_isco = kw.get("vr_isco", 2)
            self.p1 = kw.get("p1", 6.)
            self.p2 = kw.get("p2", 2.)
            self.dd = kw.get("dd", 0.2)
        elif veltype == "fromfile":
            self.radii, (self.tU1, self.tU2, self.tU3) = load_table(kw.get("file"), ["U1", "U2", "U3"])
        elif veltype == "driftframe":
            self.bfield = kw.get("bfield", None) or Bfield("bz_para")
            self.nu_parallel = kw.get("nu_parallel", 0)
        else:
            raise Exception("veltype %s not recognized in Velocity!" % veltype)

    def u_lab(self, a, r, th=np.pi / 2.):
        t = self.veltype
        if t == "zamo":
            return u_zamo(a, r)
        if t == "infall":
            return u_infall(a, r)
        if t in ("kep", "cunningham"):
            return u_kep(a, r, retrograde=self.retrograde)
        if t in ("subkep", "cunningham_subkep"):
            return u_subkep(a, r, retrograde=self.retrograde, fac_subkep=self.fac_subkep)
        if t == "general":
            return u_general(a, r, retrograde=self.retrograde, fac_subkep=self.fac_subkep, beta_phi=self.beta_phi,
                             beta_r=self.beta_r)
        if t == "gelles":
            return u_gelles(a, r, beta=self.gelles_beta, chi=self.gelles_chi)
        if t == "simfit":
            return u_grmhd_fit(a, r, ell_isco=self.ell_isco, vr_isco=self.vr_isco, p1=self.p1, p2=self.p2,
                               dd=self.dd)
        if t == "fromfile":
            return u_from_u123(a, r, self)
        return u_driftframe(a, r, bfield=self.bfield, nu_parallel=self.nu_parallel, th=th)

    def u_lab_cov(self, a, r, th=np.pi / 2.):
        return lower(metric(a, r, th), self.u_lab(a, r, th=th))

    def tetrades(self, a, r, th=np.pi / 2):
        """velocities.py:124-176."""
        m = metric(a, r, th)
        u0, u1, u2, u3 = self.u_lab(a, r, th=th)
        u0_l, u1_l, u2_l, u3_l = self.u_lab_cov(a, r, th=th)
        Nr = np.sqrt(-m["g11"] * (u0_l * u0 + u3_l * u3) * (1 + u2_l * u2))
        Nth = np.sqrt(m["g22"] * (1 + u2_l * u2))
        Nph = np.sqrt(-m["Delta"] * m["sth2"] * (u0_l * u0 + u3_l * u3))
        et = (-u0, -u1, -u2, -u3)
        ex = (u1_l * u0 / Nr, -(u0_l * u0 + u3_l * u3) / Nr, 0, u1_l * u3 / Nr)
        ey = (u2_l * u0 / Nth, u2_l * u1 / Nth, (1 + u2_l * u2) / Nth, u2_l * u3 / Nth)
        ez = (u3_l / Nph, 0, 0, -u0_l / Nph)
        return et, ex, ey, ez


def u_from_u123(a, r, v):
    """velocities.py:658-698."""
    r = _as1d(r)
    u1 = np.interp(r, v.radii, v.tU1)
    u2 = np.interp(r, v.radii, v.tU2)
    u3 = np.interp(r, v.radii, v.tU3)
    r2, a2 = r * r, a * a
    th = np.pi / 2.
    Delta = r2 - 2 * r + a2
    Sigma = r2 + a2 * np.cos(th) ** 2
    g11 = Sigma / Delta
    g22 = Sigma
    g33 = (r2 + a2 + 2 * r * (a * np.sin(th)) ** 2 / Sigma) * np.sin(th) ** 2
    gcon00 = -(r2 + a2 + 2 * r * a2 / Sigma * np.sin(th) ** 2) / Delta
    gcon03 = -2 * r * a / Sigma / Delta
    alpha = 1. / np.sqrt(-gcon00)
    gamma = np.sqrt(1. + g11 * u1 * u1 + g22 * u2 * u2 + g33 * u3 * u3)
    return gamma / alpha, u1, u2, u3 - gamma * alpha * gcon03


# ---------------------------------------------------------------------------
# emissivities (emissivities.py)
# ---------------------------------------------------------------------------

E_CHARGE, M_E, C_LIGHT, K_B = 4.80320425e-10, 9.10938356e-28, 2.99792458e10, 1.380649e-16   # emissivities.py:27-31


def emisBPL(a, r, p1=-2.0, p2=-0.5):
    """emissivities.py:121-127."""
    rh = 1 + np.sqrt(1 - a ** 2)
    return np.exp(p1 * np.log(r / rh) + p2 * np.log(r / rh) ** 2)


def emisGLM(a, r, gamma_off=-1.5, sigma=0.5, mu_ring=False):
    """emissivities.py:130-138."""
    mu = mu_ring if mu_ring else 1 - np.sqrt(1 - a ** 2)
    return np.exp(-0.5 * (gamma_off + np.arcsinh((r - mu) / sigma)) ** 2) / np.sqrt((r - mu) ** 2 + sigma ** 2)


def II_fit(x):
    """emissivities.py:141-142."""
    return 2.5651 * (1.0 + 1.92 * x ** (-1.0 / 3.0) + 0.9977 * x ** (-2.0 / 3.0)) * np.exp(-1.8899 * x ** (1.0 / 3.0))


def j_nu_thermal(ne, B, Te, nu, sin_thetaB):
    """emissivities.py:145-157."""
    Theta_e = K_B * Te / (M_E * C_LIGHT * C_LIGHT)
    nu_B = E_CHARGE * B / (2.0 * np.pi * M_E * C_LIGHT)
    nu_c = 1.5 * nu_B * Theta_e ** 2 * sin_thetaB
    x = nu / np.maximum(nu_c, 1e-40)
    return (ne * E_CHARGE ** 2 * nu) / (2 * np.sqrt(3) * C_LIGHT * (Theta_e ** 2)) * II_fit(x)


class Emissivity:
    """emissivities.py:37-118."""

    def __init__(self, emistype="bpl", **kw):
        self.emistype = emistype
        self.emiscut_in = kw.get("emiscut_in", 0)
        self.emiscut_out = kw.get("emiscut_out", 1.e10)
        if emistype == "thermal":
            self.Rb = float(kw.get("Rb", 5.0))
            self.ne0 = float(kw["ne0"])
            self.Te0 = float(kw["Te0"])
            self.B0 = float(kw["B0"])
            self.alpha_n = float(kw.get("alpha_n", 0.7))
            self.alpha_T = float(kw.get("alpha_T", 1.0))
            self.alpha_B = float(kw.get("alpha_B", 1.5))
            self.bfield_model = bool(kw.get("bfield_model", False))
        elif emistype == "bpl":
            self.p1 = kw.get("p1", -2.0)
            self.p2 = kw.get("p2", -0.5)
        elif emistype == "ring":
            self.mu_ring = kw.get("r_ring", 4.5)
            self.gamma_off = kw.get("gamma_off", 0)
            self.sigma = kw.get("sigma", 0.3)
        elif emistype == "glm":
            self.mu_ring = False
            self.gamma_off = kw.get("gamma_off", -1.5)
            self.sigma = kw.get("sigma", 0.5)
        elif emistype == "constant":
            pass
        else:
            raise Exception("emistype %s not recognized in Emissivity!" % emistype)

    def jrest(self, a, r, g=None, sinthetab=None, Bmag=None, nu_obs=230.e9, specind=0.0):
        r = np.asarray(r, dtype=float)
        if self.emistype == "thermal":
            nu_em = np.asarray(nu_obs) / np.asarray(g)
            x = r / self.Rb
            ne = self.ne0 * np.power(x, -self.alpha_n)
            Te = self.Te0 * np.power(x, -self.alpha_T)
            B = self.B0 * np.power(x, -self.alpha_B)
            if self.bfield_model:
                B = np.asarray(Bmag, dtype=float)
            return j_nu_thermal(ne, B, Te, nu_em, sinthetab)
        if self.emistype == "constant":
            j = np.ones(r.shape)
        elif self.emistype == "bpl":
            j = emisBPL(a, r, p1=self.p1, p2=self.p2)
        else:
            j = emisGLM(a, r, gamma_off=self.gamma_off, sigma=self.sigma, mu_ring=self.mu_ring)
        if g is not None:
            j = j * (g ** specind)
        if sinthetab is not None:
            j = j * (sinthetab ** (1 + specind))
        return j


# ---------------------------------------------------------------------------
# image (equatorial_images.py)
# ---------------------------------------------------------------------------

SPECIND, DISKANGLE, SINTHB, OBSFREQ = 1, 0.1, np.pi / 4., 230.e9   # equatorial_images.py:18-21


def calc_redshift(a, r, lam, eta, kr_sign, kth_sign, u0, u1, u2, u3, th=np.pi / 2):
    """equatorial_images.py:307-334, with u^theta (the main oracle's
    calc_redshift takes it too)."""
    m = metric(a, r, th)
    R = (m["r2"] + m["a2"] - a * lam) ** 2 - m["Delta"] * (eta + (lam - a) ** 2)
    TH = eta + m["a2"] * m["cth2"] - lam * lam * m["cth2"] / m["sth2"]
    return 1 / (1 * u0 - lam * u3 - np.sign(kth_sign) * u2 * np.sqrt(TH) - np.sign(kr_sign) * u1 * np.sqrt(R) / m["Delta"])


def calc_polquantities(a, r, lam, eta, kr_sign, kth_sign, velocity, bfield, th=np.pi / 2):
    """equatorial_images.py:336-487: (sinthb, kappa, pathlength, bsq)."""
    m = metric(a, r, th)
    u = velocity.u_lab(a, r, th=th)
    u0, u1, u2, u3 = u
    _, u1_l, u2_l, u3_l = lower(m, u)
    et, ex, ey, ez = velocity.tetrades(a, r, th=th)
    if bfield.fieldframe != "lab":
        # the reference's comoving branch leaves the wall normal undefined
        # (NameError at :441): no comoving field reaches the image
        raise Exception("bfield.fluidframe=%s not supported by calc_polquantities" % bfield.fieldframe)
    B1, B2, B3 = bfield.bfield_lab(a, r, th=th)
    gdet = m["Sigma"] * np.sqrt(m["sth2"])
    nrm = (0, -B2 * gdet, B1 * gdet, 0)
    b0 = B1 * u1_l + B2 * u2_l + B3 * u3_l
    b = (b0, (B1 + b0 * u1) / u0, (B2 + b0 * u2) / u0, (B3 + b0 * u3) / u0)
    bl = lower(m, b)

    def frame(e, v):
        return e[0] * v[0] + e[1] * v[1] + e[2] * v[2] + e[3] * v[3]

    Bp = [frame(e, bl) for e in (ex, ey, ez)]
    nv = [frame(e, nrm) for e in (ex, ey, ez)]
    Bp_mag = np.sqrt(Bp[0] ** 2 + Bp[1] ** 2 + Bp[2] ** 2)
    bsq = Bp_mag ** 2
    nmag = np.sqrt(nv[0] ** 2 + nv[1] ** 2 + nv[2] ** 2)
    nu = [x / nmag for x in nv]
    R = (m["r2"] + m["a2"] - a * lam) ** 2 - m["Delta"] * (eta + (lam - a) ** 2)
    TH = eta + m["a2"] * m["cth2"] - lam * lam * m["cth2"] / m["sth2"]
    kl = (-1, kr_sign * np.sqrt(R) / m["Delta"], kth_sign * np.sqrt(TH), lam)
    k0 = m["gu00"] * kl[0] + m["gu03"] * kl[3]
    k1 = m["gu11"] * kl[1]
    k2 = m["gu22"] * kl[2]
    k3 = m["gu33"] * kl[3] + m["gu03"] * kl[0]
    kp = [frame(e, kl) for e in (ex, ey, ez)]
    kp_mag = np.sqrt(kp[0] ** 2 + kp[1] ** 2 + kp[2] ** 2)
    kdotn = kp[0] * nu[0] + kp[1] * nu[1] + kp[2] * nu[2]
    pathlength = np.abs(kp_mag / kdotn)
    fx = (kp[1] * Bp[2] - kp[2] * Bp[1]) / kp_mag
    fy = (kp[2] * Bp[0] - kp[0] * Bp[2]) / kp_mag
    fz = (kp[0] * Bp[1] - kp[1] * Bp[0]) / kp_mag
    sinthb = np.sqrt(fx ** 2 + fy ** 2 + fz ** 2) / Bp_mag
    f = [ex[i] * fx + ey[i] * fy + ez[i] * fz for i in range(4)]
    A = k0 * f[1] - k1 * f[0] + a * m["sth2"] * (k1 * f[3] - k3 * f[1])
    B = ((m["r2"] + m["a2"]) * (k3 * f[2] - k2 * f[3]) - a * (k0 * f[2] - k2 * f[0])) * np.sin(th)
    kappa = (r - 1j * a * np.cos(th)) * (A - 1j * B)
    return sinthb, kappa, pathlength, bsq


def calc_pathlength_equatorial(a, r, lam, eta, kr_sign, kth_sign, u0, u1, u2, u3, th=np.pi / 2, diskangle=DISKANGLE):
    """equatorial_images.py:489-522."""
    gg = calc_redshift(a, r, lam, eta, kr_sign, kth_sign, u0, u1, u2, u3, th=th)
    if (isinstance(u2, np.ndarray) and np.any(u2 != 0)) or (not isinstance(u2, np.ndarray) and u2 != 0):
        raise Exception("calc_pathlength assumes that u2==0!")
    TH = eta + a * a * np.cos(th) ** 2 - lam * lam * np.cos(th) ** 2 / np.sin(th) ** 2
    kyhat = -(kth_sign * np.sqrt(TH)) / r
    return diskangle * r * ((1 / gg) / kyhat)


def calc_evpa(a, th_o, alpha, beta, kappa):
    """equatorial_images.py:524-544."""
    mu = -(alpha + a * np.sin(th_o))
    k1, k2 = np.real(kappa), np.imag(kappa)
    den = (beta ** 2 + mu ** 2) * (k1 ** 2 + k2 ** 2)
    cos2chi = ((beta * k1 + mu * k2) ** 2 - (mu * k1 - beta * k2) ** 2) / den
    sin2chi = (2 * (beta * k1 + mu * k2) * (mu * k1 - beta * k2)) / den
    return cos2chi, sin2chi


def Iobs(a, r_o, th_o, mbar, alpha, beta, emissivity, velocity, bfield, polarization=False, pathlength=False,
         specind=SPECIND, diskangle=DISKANGLE, nu_obs=OBSFREQ, th_s=np.pi / 2):
    """equatorial_images.py:124-264: (I, Q, U, g, r_s, sin_thb, Ir, Imax,
    Nmax[, lp, Ie])."""
    alpha, beta = _as1d(alpha), _as1d(beta)
    rh = 1 + np.sqrt(1 - a ** 2)
    lam, eta = conserved(a, th_o, alpha, beta)
    r_s, Ir, Imax, Nmax = r_equatorial(a, r_o, th_o, mbar, alpha, beta)
    n = alpha.shape
    g, sin_thb, I, Q, U = (np.zeros(n) for _ in range(5))
    lp, Ie = np.zeros(n), np.zeros(n)
    zm = (Nmax < mbar) + (Nmax == -1) + (r_s <= rh)
    if emissivity.emiscut_in > 0:
        zm = zm + (r_s < emissivity.emiscut_in)
    if emissivity.emiscut_out > 0:
        zm = zm + (r_s > emissivity.emiscut_out)
    k = ~zm
    if np.any(k):
        krs = radial_momentum_sign(a, th_o, alpha[k], beta[k], Ir[k], Imax[k])
        kths = theta_momentum_sign(th_o, mbar)
        u0, u1, u2, u3 = velocity.u_lab(a, r_s[k], th=th_s)
        gg = calc_redshift(a, r_s[k], lam[k], eta[k], krs, kths, u0, u1, u2, u3, th=th_s)
        g[k] = gg
        if polarization:
            sinthb, kappa, _, bsq = calc_polquantities(a, r_s[k], lam[k], eta[k], krs, kths, velocity, bfield, th=th_s)
            cos2chi, sin2chi = calc_evpa(a, th_o, alpha[k], beta[k], kappa)
        else:
            sinthb = SINTHB
            bsq = bfield.bsq(a, r_s[k], velocity, th=th_s)
        sin_thb[k] = sinthb
        Bmag = np.sqrt(bsq)
        if emissivity.emistype == "thermal":
            Bmag = (Bmag / np.sqrt(bfield.bsq(a, emissivity.Rb, velocity, th=th_s))) * emissivity.B0
        Iemis = emissivity.jrest(a, r_s[k], gg, sinthb, Bmag=Bmag, nu_obs=nu_obs, specind=specind)
        if pathlength:
            llp = calc_pathlength_equatorial(a, r_s[k], lam[k], eta[k], krs, kths, u0, u1, u2, u3, th=th_s,
                                             diskangle=diskangle)
            Iemis = Iemis * np.abs(llp)
            I[k] = gg ** 3 * Iemis
            Ie[k] = Iemis
            lp[k] = llp
        else:
            I[k] = gg ** 2 * Iemis
        if polarization:
            Q[k] = cos2chi * I[k]
            U[k] = sin2chi * I[k]
    out = (I, Q, U, g, r_s, sin_thb, Ir, Imax, Nmax)
    return out + (lp, Ie) if pathlength else out


def pixel_grid(alpha_min, alpha_max, beta_min, beta_max, psize):
    """make_image's screen grid (equatorial_images.py:49-57)."""
    n_alpha = int(np.floor(alpha_max - alpha_min) / psize)
    alphas = np.linspace(alpha_min, alpha_min + n_alpha * psize, n_alpha)
    n_beta = int(np.floor(beta_max - beta_min) / psize)
    betas = np.linspace(beta_min, beta_min + n_beta * psize, n_beta)
    A, B = np.meshgrid(alphas, betas)
    return A.flatten(), B.flatten()


def make_image_pixels(a, r_o, th_o, mbar_max, alpha, beta, nmax_only=False, emissivity=None, bfield=None,
                      velocity=None, polarization=False, pathlength=False, specind=SPECIND, diskangle=DISKANGLE,
                      nu_obs=OBSFREQ):
    """make_image (equatorial_images.py:23-122) on given pixels: (I, Q, U,
    r, t, g, sinthb, n, np[, lp, Ie])."""
    emissivity = emissivity or Emissivity("bpl")
    bfield = bfield or Bfield("rad")
    velocity = velocity or Velocity("zamo")
    alpha, beta = _as1d(alpha), _as1d(beta)
    ns = mbar_max + 1
    shape = (len(alpha), ns)
    I, Q, U, R, T, G, S = (np.zeros(shape) for _ in range(7))
    lp, Ie = np.zeros(shape), np.zeros(shape)
    lam, eta = conserved(a, th_o, alpha, beta)
    from .kgeo_oracle import mino_total, radial_roots
    r1, r2, r3, r4, _ = radial_roots(a, lam, eta)
    Imax = mino_total(a, r_o, eta, r1, r2, r3, r4)
    n_pol = n_poloidal_orbits(a, th_o, alpha, beta, Imax)
    from .kgeo_oracle import n_equatorial_crossings
    if nmax_only:
        n = n_equatorial_crossings(a, th_o, alpha, beta, Imax)
        return (I, Q, U, R, T, G, S, n, n_pol)
    n = None
    for mbar in range(ns):
        d = Iobs(a, r_o, th_o, mbar, alpha, beta, emissivity, velocity, bfield, polarization=polarization,
                 pathlength=pathlength, specind=specind, diskangle=diskangle, nu_obs=nu_obs)
        I[:, mbar], Q[:, mbar], U[:, mbar], G[:, mbar], R[:, mbar], S[:, mbar], T[:, mbar] = d[:7]
        n = d[8]
        if pathlength:
            lp[:, mbar], Ie[:, mbar] = d[9], d[10]
    out = (I, Q, U, R, T, G, S, n, n_pol)
    return out + (lp, Ie) if pathlength else out
