// Emission models of the equatorial disk and the polarized observed
// intensity (the reference's Iobs): four-velocities (velocities.py), lab-frame
// magnetic fields (bfields.py), rest-frame emissivities (emissivities.py),
// the comoving-frame polarization and the Penrose-Walker transport to the
// screen (equatorial_images.py).  Host-compilable (KG_INLINE), so the CPU
// test build evaluates the code the kernels run.
//
// Every model is evaluated in the reference's arithmetic order (a**2 as a*a,
// r**1.5 as pow, the metric as equatorial_images.py:351-369 writes it).  The
// equatorial plane is theta = np.pi/2 with numpy's cos and sin of it
// (6.123233995736766e-17 and 1), which enter the metric and the field
// formulas exactly as in the reference.

#pragma once

#include "../../include/kgeo.h"
#include "kg_counts.hpp"
#include "kg_equatorial.hpp"

namespace kg {

constexpr double kPiD = 3.141592653589793;
constexpr double kHalfPiD = 1.5707963267948966;
constexpr double kNpCosHalfPi = 6.123233995736766e-17;  // np.cos(np.pi/2)

// A polar angle with the cos and sin numpy gives for it.
struct Angle {
  double th, cth, sth;
};

KG_INLINE Angle equator() { return Angle{kHalfPiD, kNpCosHalfPi, 1.0}; }

KG_INLINE Angle angle_of(double th) {
  if (th == kHalfPiD) return equator();
  return Angle{th, kcos(th), ksin(th)};
}

KG_INLINE double kasinh(double x) { return ::asinh(x); }

KG_INLINE double np_sign(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : x); }

// Boyer-Lindquist metric at (r, th) (equatorial_images.py:351-369).
struct Metric {
  double r2, a2, cth2, sth2, Delta, Sigma;
  double g00, g11, g22, g33, g03;
  double gu00, gu11, gu22, gu33, gu03;
};

KG_INLINE Metric metric_at(double a, double r, const Angle& an) {
  Metric m;
  m.a2 = a * a;
  m.r2 = r * r;
  m.cth2 = an.cth * an.cth;
  m.sth2 = an.sth * an.sth;
  m.Delta = m.r2 - 2.0 * r + m.a2;
  m.Sigma = m.r2 + m.a2 * m.cth2;
  m.g00 = -(1.0 - 2.0 * r / m.Sigma);
  m.g11 = m.Sigma / m.Delta;
  m.g22 = m.Sigma;
  m.g33 = (m.r2 + m.a2 + 2.0 * r * m.a2 * m.sth2 / m.Sigma) * m.sth2;
  m.g03 = -2.0 * r * a * m.sth2 / m.Sigma;
  m.gu00 = -(m.r2 + m.a2 + 2.0 * r * m.a2 * m.sth2 / m.Sigma) / m.Delta;
  m.gu11 = m.Delta / m.Sigma;
  m.gu22 = 1.0 / m.Sigma;
  m.gu33 = (m.Delta - m.a2 * m.sth2) / (m.Sigma * m.Delta * m.sth2);
  m.gu03 = -(2.0 * a * r) / (m.Sigma * m.Delta);
  return m;
}

struct V4 {
  double v[4];
};

KG_INLINE V4 lower(const Metric& m, const V4& u) {
  return V4{{m.g00 * u.v[0] + m.g03 * u.v[3], m.g11 * u.v[1], m.g22 * u.v[2], m.g33 * u.v[3] + m.g03 * u.v[0]}};
}

KG_INLINE double contract(const V4& a, const V4& b) {
  return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2] + a.v[3] * b.v[3];
}

// np.interp over a 4 x n column-block table: column col at x (clamped to the
// end values outside [r_0, r_{n-1}]).
KG_INLINE double table_interp(const double* t, int64_t n, int col, double x) {
  const double* xp = t;
  const double* fp = t + col * n;
  if (!(x >= xp[0])) return (x < xp[0]) ? fp[0] : x;  // NaN stays NaN
  if (x >= xp[n - 1]) return fp[n - 1];
  int64_t lo = 0, hi = n - 1;  // xp[lo] <= x < xp[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid;
    else hi = mid;
  }
  if (xp[lo] == x) return fp[lo];
  const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
  return slope * (x - xp[lo]) + fp[lo];
}

// ---------------------------------------------------------------------------
// special functions for the field models
// ---------------------------------------------------------------------------

// Dilogarithm Li2(x) for real x <= 1: the Bernoulli series in
// u = -ln(1 - x) on [-1, 1/2] ('t Hooft and Veltman 1979), with the
// reflection Li2(x) = pi^2/6 - ln x ln(1-x) - Li2(1-x) above 1/2 and the
// inversion Li2(x) = -pi^2/6 - ln^2(-x)/2 - Li2(1/x) below -1.
KG_INLINE double li2_core(double x) {  // -1 <= x <= 1/2
  const double u = -klog1p(-x);
  const double u2 = u * u;
  // B_n / (n + 1)! for n = 2, 4, ..., 20 (odd powers u^3 ... u^21)
  constexpr double c[10] = {0.027777777777777776,  -0.0002777777777777778,  4.72411186696901e-06,
                            -9.185773074661964e-08, 1.8978869988971e-09,    -4.0647616451442256e-11,
                            8.921691020456452e-13,  -1.9939295860721074e-14, 4.518980029619918e-16,
                            -1.0356517612181247e-17};
  double s = c[9];
  for (int k = 8; k >= 0; --k) s = s * u2 + c[k];
  return u - 0.25 * u2 + u * u2 * s;
}

KG_INLINE double li2(double x) {
  constexpr double pi2_6 = kPiD * kPiD / 6.0;
  if (x == 1.0) return pi2_6;
  if (x > 0.5) return pi2_6 - klog(x) * klog1p(-x) - li2_core(1.0 - x);
  if (x < -1.0) {
    const double l = klog(-x);
    return -pi2_6 - 0.5 * l * l - li2_core(1.0 / x);
  }
  return li2_core(x);
}

// Real part of Li2(x) for x > 1: pi^2/3 - ln^2(x)/2 - Li2(1/x).
KG_INLINE double re_li2_above1(double x) {
  const double l = klog(x);
  return kPiD * kPiD / 3.0 - 0.5 * l * l - li2(1.0 / x);
}

// The BZ monopole radial function f(r) of bfields.py:7-13 (real part) and its
// derivative.  For r < 4 the closed form; for r >= 4 its expansion in
// x = 2/r, f = sum_{m>=1} a_m x^m - ln x sum_{m>=2} b_m x^m with
// a_m = 2/(m+3)^2 - 3/(2(m+2)^2) (+ 1/6 for m = 1), b_m = 2/(m+3) -
// 3/(2(m+2)): the closed form's terms of order x^-2 .. x^0 cancel, leaving
// f = x/8 + O(x^2 ln x) = 1/(4r) + ..., which the series sums without the
// closed form's cancellation.  Beyond r = 100 the
// reference takes f = 1/(4r), f' = -1/(4r^2) (bfields.py:560-563).
KG_INLINE void bz_f(double r, double& f, double& fp) {
  if (r > 100.0) {
    f = 1.0 / (4.0 * r);
    fp = -1.0 / (4.0 * (r * r));
    return;
  }
  if (r >= 4.0) {
    const double x = 2.0 / r;
    const double lx = klog(x);
    double sa = 0.0, sb = 0.0, dsa = 0.0, dsb = 0.0, dsb0 = 0.0;
    double xm1 = 1.0;  // x^(m-1)
    for (int m = 1; m <= 64; ++m) {
      // (1/(3r) = x/6 of the closed form's remainder joins the x^1 term)
      const double am = 2.0 / ((m + 3.0) * (m + 3.0)) - 1.5 / ((m + 2.0) * (m + 2.0)) + ((m == 1) ? 1.0 / 6.0 : 0.0);
      const double bm = 2.0 / (m + 3.0) - 1.5 / (m + 2.0);
      const double xm = xm1 * x;
      sa += am * xm;
      sb += bm * xm;
      dsa += m * am * xm1;
      dsb += m * bm * xm1;
      dsb0 += bm * xm1;
      xm1 = xm;
      if (xm < 1e-18) break;
    }
    f = sa - lx * sb;
    const double dfdx = dsa - dsb0 - lx * dsb;
    fp = dfdx * (-0.5 * x * x);
    return;
  }
  const double x = 2.0 / r;
  const double y = 0.5 * (r - 2.0);  // r/2 - 1
  const double L = klog1p(y);        // ln(r/2)
  double A;
  if (r == 2.0) {
    A = kPiD * kPiD / 6.0;  // the limit: ln|1 - x| ln(r/2) -> 0
  } else {
    const double rl = (x > 1.0) ? re_li2_above1(x) : li2(x);
    A = rl - klog(kfabs(1.0 - x)) * L;
  }
  const double B = r * r * (2.0 * r - 3.0) / 8.0;
  const double C = (1.0 + 3.0 * r - 6.0 * r * r) * L / 12.0 + 11.0 / 72.0 + 1.0 / (3.0 * r) + r / 2.0 - r * r / 2.0;
  f = A * B + C;
  const double l1py = (y == 0.0) ? 1.0 : L / y;  // ln(r/2) / (r/2 - 1)
  const double dA = -l1py / r;                   // -2 ln(r/2) / (r (r - 2))
  const double dB = 3.0 * r * (r - 1.0) / 4.0;
  const double dC = (3.0 - 12.0 * r) * L / 12.0 + (1.0 + 3.0 * r - 6.0 * r * r) / (12.0 * r) - 1.0 / (3.0 * r * r) +
                    0.5 - r;
  fp = dA * B + A * dB + dC;
}

// Principal branch of the Lambert W function for real x >= -1/e: Halley's
// iteration on w e^w = x from a branch-point expansion (x near -1/e),
// Winitzki's log1p guess (moderate x) or the asymptotic L1 - L2 + L2/L1.
KG_INLINE double lambert_w0(double x) {
  constexpr double e = 2.718281828459045;
  if (x == 0.0) return 0.0;
  double w;
  if (x < -0.25) {
    const double q = 2.0 * (e * x + 1.0);
    if (q <= 0.0) return -1.0;
    const double p = ksqrt(q);
    w = -1.0 + p - p * p / 3.0 + 11.0 / 72.0 * p * p * p;
  } else if (x < 3.0) {
    const double l = klog1p(x);
    w = l * (1.0 - klog1p(l) / (2.0 + l));
  } else {
    const double l1 = klog(x), l2 = klog(l1);
    w = l1 - l2 + l2 / l1;
  }
  for (int it = 0; it < 12; ++it) {
    const double ew = kexp(w);
    const double fw = w * ew - x;
    const double wp1 = w + 1.0;
    if (wp1 == 0.0) break;
    const double dw = fw / (ew * wp1 - (w + 2.0) * fw / (2.0 * wp1));
    w -= dw;
    if (!(kfabs(dw) > 1e-16 * (1.0 + kfabs(w)))) break;
  }
  return w;
}

// ---------------------------------------------------------------------------
// lab-frame magnetic fields (bfields.py)
// ---------------------------------------------------------------------------

struct BLab {
  double B1, B2, B3, omega;
};

KG_INLINE bool bfield_has_omega(int kind) {
  return kind == KGEO_BF_MONOPOLE_A || kind == KGEO_BF_BZ_MONOPOLE || kind == KGEO_BF_BZ_GUESS ||
         kind == KGEO_BF_BZ_PARA || kind == KGEO_BF_POWER || kind == KGEO_BF_GEN_POWER;
}

// Bfield_simple / Bfield_simple_rm1 (bfields.py:429-477), equatorial only.
KG_INLINE BLab field_simple(double a, double r, double cr, double cv, double cp, bool rm1) {
  const Angle q = equator();
  const double Sigma = r * r + (a * a) * (q.cth * q.cth);
  const double gdet = kfabs(q.sth) * Sigma;
  const double rad = rm1 ? (r * q.sth / gdet) : (q.sth / gdet);
  BLab b;
  b.B1 = cr * rad + cv * (r * q.cth / gdet);
  b.B2 = cv * (-q.sth / gdet);
  b.B3 = cp * (1.0 / gdet);
  b.omega = NAN;
  return b;
}

// Bfield_monopoleA (bfields.py:480-501)
KG_INLINE BLab field_monopole_a(double a, double r, const Angle& an, double C, double omegafac) {
  const double rh = 1.0 + ksqrt(1.0 - a * a);
  const double omegah = a / (2.0 * rh);
  const double phih = 4.0 * kPiD * C;
  const double Delta = r * r - 2.0 * r + a * a;
  const double Sigma = r * r + (a * a) * (an.cth * an.cth);
  const double omegaf = omegafac * omegah;
  const double s2 = an.sth * an.sth;
  const double I = phih * s2 * rh * (omegaf - omegah) / (rh * rh + (a * a) * (an.cth * an.cth));
  BLab b;
  b.B1 = (phih / (4.0 * kPiD)) / Sigma;
  b.B2 = 0.0;
  b.B3 = I / (2.0 * kPiD * Delta * s2);
  b.omega = omegaf;
  return b;
}

// Bfield_BZmagic, 'bz_guess' (bfields.py:504-542)
KG_INLINE BLab field_bz_guess(double a, double r, const Angle& an, double C) {
  const double a2 = a * a, r2 = r * r;
  const double sth2 = an.sth * an.sth;
  const double Delta = r2 - 2.0 * r + a2;
  const double Pi = (r2 + a2) * (r2 + a2) - a2 * Delta * sth2;
  const double rH = 1.0 + ksqrt(1.0 - a2);
  const double OmegaH = a / (2.0 * rH);
  const double spin = kfabs(a);
  const double f2 = (6.0 * kPiD * kPiD - 49.0) / 72.0;
  const double omega2 = 1.0 / 32.0 - (4.0 * f2 - 1.0) * sth2 / 64.0;
  double OmegaBZ = spin * (1.0 / 8.0) + (spin * spin * spin) * omega2;
  if (a < 0.0) OmegaBZ = -OmegaBZ;
  BLab b;
  b.B1 = C / r2;
  b.B2 = 0.0 * b.B1;
  b.B3 = (-1.0 * (OmegaH - OmegaBZ) * ksqrt(Pi) / Delta) * b.B1;
  b.omega = OmegaBZ;
  return b;
}

// Bfield_BZmonopole (bfields.py:545-592) with f(r), f'(r) from bz_f.
KG_INLINE BLab field_bz_monopole(double a, double r, const Angle& an, double C, bool second_order) {
  const double a2 = a * a, r2 = r * r;
  const double sth = an.sth, cth = an.cth;
  const double cth2 = cth * cth, sth2 = sth * sth;
  const double Delta = r2 - 2.0 * r + a2;
  const double Sigma = r2 + a2 * cth2;
  const double gdet = sth * Sigma;
  double fr, frp;
  bz_f(r, fr, frp);
  const double cos2th = kcos(2.0 * an.th);
  const double f2 = (6.0 * kPiD * kPiD - 49.0) / 72.0;
  const double omega2 = 1.0 / 32.0 - sth2 * (4.0 * f2 - 1.0) / 64.0;
  BLab b;
  if (second_order) {
    const double d = r2 - 2.0 * r;
    b.B1 = C / r2 + C * a2 * (-cth2 / (r2 * r2) + fr * (0.5 + 1.5 * cos2th) / r2);
    b.B2 = -C * a2 * frp * cth * sth / r2;
    b.B3 = -C * 0.125 * a / d + C * (a * a * a) * (0.125 / (d * d) - (omega2 + 0.25 * cth2 * fr) / d);
  } else {
    const double dphidtheta = C * sth * (1.0 + 0.5 * a2 * fr * (1.0 + 3.0 * cos2th));
    const double dphidr = C * a2 * sth2 * cth * frp;
    const double I = -1.0 * C * 2.0 * kPiD * sth2 * (a / 8.0 + (a * a * a) * (omega2 + 0.25 * fr * cth2));
    b.B1 = dphidtheta / gdet;
    b.B2 = -dphidr / gdet;
    b.B3 = I / (2.0 * kPiD * Delta * sth2);
  }
  const double spin = kfabs(a);
  b.omega = spin / 8.0 + (spin * spin * spin) * omega2;
  if (a < 0.0) b.omega = -b.omega;
  return b;
}

// Bfield_BZpara (bfields.py:609-673): Omega from the principal Lambert W,
// zero where its argument is below -1/e.
KG_INLINE BLab field_bz_para(double a, double r, const Angle& an, double C) {
  const double a2 = a * a, r2 = r * r;
  const double rp = 1.0 + ksqrt(1.0 - a2);
  const double sth = an.sth, cth = an.cth;
  const double abscth = kfabs(cth);
  const double Delta = r2 - 2.0 * r + a2;
  const double Sigma = r2 + a2 * (cth * cth);
  const double gdet = sth * Sigma;
  const double ln2 = 0.6931471805599453, ln4 = 1.3862943611198906;
  const double psi = r * (1.0 - abscth) + rp * (1.0 + abscth) * (1.0 - klog(1.0 + abscth)) - 2.0 * rp * (1.0 - ln2);
  const double dpsidtheta = np_sign(cth) * sth * (r + rp * klog(1.0 + abscth));
  const double dpsidr = 1.0 - abscth;
  const double arg = -psi / rp + ln4;
  double om = 0.0;
  if (arg >= -1.0 / 2.718281828459045) {
    const double w = lambert_w0(arg);
    const double xf = kexp(w);
    const double yf = 1.0 + psi / (1.0 + w) / (xf * (2.0 - xf));
    om = a / 4.0 / yf;
  }
  BLab b;
  b.omega = C * om;
  const double I = -4.0 * kPiD * psi * b.omega * np_sign(cth);
  b.B1 = C * dpsidtheta / gdet;
  b.B2 = -C * dpsidr / gdet;
  b.B3 = I / (2.0 * kPiD * Delta * (sth * sth));
  return b;
}

// Bfield_power (bfields.py:618-624, 676-709)
KG_INLINE BLab field_power(double a, double r, const Angle& an, double p, double C, bool usemono) {
  const double a2 = a * a, r2 = r * r;
  const double sth = an.sth, cth = an.cth;
  const double abscth = kfabs(cth);
  const double Delta = r2 - 2.0 * r + a2;
  const double Sigma = r2 + a2 * (cth * cth);
  const double gdet = sth * Sigma;
  const double rpow = kpow(r, p);
  const double psi = rpow * (1.0 - abscth);
  const double dpsidtheta = np_sign(cth) * sth * rpow;
  const double dpsidr = p * psi / r;
  double om;
  if (usemono || p == 0.0) {
    om = a / 8.0;
  } else {
    const double rp = 1.0 + ksqrt(1.0 - a2);
    const double ch = 1.0 - psi * kpow(rp, -p);
    om = a / (4.0 + 8.0 / (1.0 + ch));
  }
  const double I = (p > 0.0) ? -4.0 * kPiD * psi * om * np_sign(cth) : -2.0 * kPiD * psi * (2.0 - psi) * om * np_sign(cth);
  BLab b;
  b.B1 = C * dpsidtheta / gdet;
  b.B2 = -C * dpsidr / gdet;
  b.B3 = I / (2.0 * kPiD * Delta * (sth * sth));
  b.omega = om;
  return b;
}

// Bfield_gen_power (bfields.py:712-746)
KG_INLINE BLab field_gen_power(double a, double r, const Angle& an, double n_I, double p_val, bool above) {
  const double a2 = a * a, r2 = r * r;
  const double rp = 1.0 + ksqrt(1.0 - a2);
  const double sth = an.sth, cth = an.cth;
  const double Delta = r2 - 2.0 * r + a2;
  const double Sigma = r2 + a2 * (cth * cth);
  const double gdet = sth * Sigma;
  const double sgn = above ? 1.0 : -1.0;
  const double I0 = sgn * (a * kPiD * (sth * sth)) / (rp * rp + a2 * (cth * cth));
  const double dpsidtheta = sgn * sth * kpow(r / rp, p_val);
  const double I = I0 * kpow(r / rp, -1.0 * n_I);
  const double dpsidr = (1.0 - kfabs(cth)) * ((p_val * kpow(r, p_val - 1.0)) / kpow(rp, p_val));
  BLab b;
  b.B1 = dpsidtheta / gdet;
  b.B2 = -1.0 * (dpsidr / gdet);
  b.B3 = I / (2.0 * kPiD * Delta * (sth * sth));
  b.omega = 0.5 * (a / (2.0 * rp));
  return b;
}

// Bfield.bfield_lab and the field-line angular velocity (bfields.py:93-131,
// :175-196); omega is NaN for kinds without one.
KG_INLINE BLab bfield_lab(const kgeo_bfield_model& f, double a, double r, const Angle& an) {
  switch (f.kind) {
    case KGEO_BF_SIMPLE: return field_simple(a, r, f.Cr, f.Cvert, f.Cph, false);
    case KGEO_BF_SIMPLE_RM1: return field_simple(a, r, f.Cr, f.Cvert, f.Cph, true);
    case KGEO_BF_MONOPOLE_A: return field_monopole_a(a, r, an, f.C, f.omegafac);
    case KGEO_BF_BZ_MONOPOLE: return field_bz_monopole(a, r, an, f.C, f.secondorder_only != 0);
    case KGEO_BF_BZ_GUESS: return field_bz_guess(a, r, an, f.C);
    case KGEO_BF_BZ_PARA: return field_bz_para(a, r, an, f.C);
    case KGEO_BF_POWER: return field_power(a, r, an, f.p, f.C, f.usemono != 0);
    case KGEO_BF_GEN_POWER: return field_gen_power(a, r, an, f.n_I, f.p_val, f.is_above != 0);
    default: {  // KGEO_BF_FROMFILE (bfields.py:790-806)
      BLab b;
      b.B1 = table_interp(f.table, f.table_n, 1, r);
      b.B2 = table_interp(f.table, f.table_n, 2, r);
      b.B3 = table_interp(f.table, f.table_n, 3, r);
      b.omega = NAN;
      return b;
    }
  }
}

// ---------------------------------------------------------------------------
// four-velocities (velocities.py)
// ---------------------------------------------------------------------------

KG_INLINE double isco_radius(double a, double s) {
  const double z1 = 1.0 + kcbrt(1.0 - a * a) * (kcbrt(1.0 + a) + kcbrt(1.0 - a));
  const double z2 = ksqrt(3.0 * (a * a) + z1 * z1);
  return 3.0 + z2 - s * ksqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

// u_zamo (velocities.py:179-208), equatorial only
KG_INLINE V4 u_zamo(double a, double r) {
  const Angle q = equator();
  const double Sigma = r * r + (a * a) * (q.cth * q.cth);
  const double g00 = -(1.0 - 2.0 * r / Sigma);
  const double as = a * q.sth;
  const double g33 = (r * r + a * a + 2.0 * r * (as * as) / Sigma) * (q.sth * q.sth);
  const double g03 = -2.0 * r * a * (q.sth * q.sth) / Sigma;
  const double v3 = -g03 / g33;
  const double u0 = ksqrt(-1.0 / (g00 + 2.0 * g03 * v3 + g33 * v3 * v3));
  return V4{{u0, 0.0, 0.0, u0 * v3}};
}

// u_infall (velocities.py:211-226)
KG_INLINE V4 u_infall(double a, double r) {
  const double Delta = r * r + a * a - 2.0 * r;
  const double Xi = (r * r + a * a) * (r * r + a * a) - Delta * (a * a);
  return V4{{Xi / (r * r * Delta), -ksqrt(2.0 * r * (r * r + a * a)) / (r * r), 0.0, 2.0 * a / (r * Delta)}};
}

// u_kep (velocities.py:229-281), either direction
KG_INLINE V4 u_kepler(double a, double r, double s) {
  const double ri = isco_radius(a, s);
  const double spin = kfabs(a);
  const double asign = np_sign(a);
  if (r >= ri) {
    const double r15 = kpow(r, 1.5);
    const double Om = asign * s / (r15 + s * spin);
    const double u0 = (r15 + s * spin) / ksqrt(kpow(r, 3.0) - 3.0 * (r * r) + 2.0 * s * spin * r15);
    return V4{{u0, 0.0, 0.0, Om * u0}};
  }
  const double Delta = r * r - 2.0 * r + a * a;
  const double sri = ksqrt(ri);
  const double ell = s * asign * (ri * ri + a * a - s * 2.0 * spin * sri) / (kpow(ri, 1.5) - 2.0 * sri + s * spin);
  const double gam = ksqrt(1.0 - 2.0 / (3.0 * ri));
  const double H = (2.0 * r - a * ell) / Delta;
  return V4{{gam * (1.0 + (2.0 / r) * (1.0 + H)), -ksqrt(2.0 / (3.0 * ri)) * kpow(ri / r - 1.0, 1.5), 0.0,
             gam * (ell + a * H) / (r * r)}};
}

// u_subkep (velocities.py:284-359)
KG_INLINE V4 u_subkep(double a, double r, double fac, double s) {
  const double ri = isco_radius(a, s);
  const double spin = kfabs(a);
  const double asign = np_sign(a);
  auto ell_gam = [&](double rr, double& ell, double& gam) {
    const double D = rr * rr - 2.0 * rr + a * a;
    const double X = (rr * rr + a * a) * (rr * rr + a * a) - D * (a * a);
    const double srr = ksqrt(rr);
    ell = asign * s * (rr * rr + spin * spin - s * 2.0 * spin * srr) / (kpow(rr, 1.5) - 2.0 * srr + s * spin);
    ell = ell * fac;
    gam = ksqrt(D / (X / (rr * rr) - 4.0 * a * ell / rr - (1.0 - 2.0 / rr) * (ell * ell)));
  };
  const double Delta = r * r - 2.0 * r + a * a;
  if (r >= ri) {
    double ell, gam;
    ell_gam(r, ell, gam);
    const double H = (2.0 * r - a * ell) / Delta;
    const double chi = 1.0 / (1.0 + (2.0 / r) * (1.0 + H));
    const double Om = (chi / (r * r)) * (ell + a * H);
    return V4{{gam / chi, 0.0, 0.0, (gam / chi) * Om}};
  }
  const double Xi = (r * r + a * a) * (r * r + a * a) - Delta * (a * a);
  double ell_i, gam_i;
  ell_gam(ri, ell_i, gam_i);
  const double H = (2.0 * r - a * ell_i) / Delta;
  const double chi = 1.0 / (1.0 + (2.0 / r) * (1.0 + H));
  const double Om = (chi / (r * r)) * (ell_i + a * H);
  const double nu = (r / Delta) * ksqrt(Xi / (r * r) - 4.0 * a * ell_i / r - (1.0 - 2.0 / r) * (ell_i * ell_i) -
                                        Delta / (gam_i * gam_i));
  return V4{{gam_i / chi, -gam_i * (Delta / (r * r)) * nu, 0.0, (gam_i / chi) * Om}};
}

// u_general (velocities.py:442-471)
KG_INLINE V4 u_general(double a, double r, double fac, double beta_phi, double beta_r, double s) {
  const double Delta = r * r + a * a - 2.0 * r;
  const V4 ui = u_infall(a, r);
  const V4 us = u_subkep(a, r, fac, s);
  const double Om_i = ui.v[3] / ui.v[0];
  const double Om_s = us.v[3] / us.v[0];
  const double u1 = us.v[1] + (1.0 - beta_r) * (ui.v[1] - us.v[1]);
  const double Om = Om_s + (1.0 - beta_phi) * (Om_i - Om_s);
  double u0 = ksqrt(1.0 + (r * r) * (u1 * u1) / Delta);
  const double oma = 1.0 - a * Om;
  u0 = u0 / ksqrt(1.0 - (r * r + a * a) * (Om * Om) - (2.0 / r) * (oma * oma));
  return V4{{u0, u1, 0.0, u0 * Om}};
}

// u_gelles (velocities.py:362-389), equatorial only
KG_INLINE V4 u_gelles(double a, double r, double beta, double chi) {
  const Angle q = equator();
  const double a2 = a * a, r2 = r * r;
  const double sth2 = q.sth * q.sth;
  const double Delta = r2 - 2.0 * r + a2;
  const double Xi = (r2 + a2) * (r2 + a2) - Delta * a2 * sth2;
  const double omegaz = 2.0 * a * r / Xi;
  const double gamma = 1.0 / ksqrt(1.0 - beta * beta);
  const double sXD = ksqrt(Xi / Delta);
  return V4{{(gamma / r) * sXD, (beta * gamma * kcos(chi) / r) * ksqrt(Delta), 0.0,
             (gamma * omegaz / r) * sXD + (r * beta * gamma * ksin(chi)) / ksqrt(Xi)}};
}

// u_grmhd_fit, 'simfit' (velocities.py:392-439)
KG_INLINE V4 u_simfit(double a, double r, double ell_isco, double vr_isco, double p1, double p2, double dd) {
  const double r_isco = isco_radius(a, 1.0);
  const Metric m = metric_at(a, r, equator());
  const double x = r / r_isco;
  const double ell = ell_isco * kpow(x, 0.5);
  const double vr = -vr_isco * kpow(x, -p1) * kpow(0.5 * (1.0 + kpow(x, 1.0 / dd)), (p1 - p2) * dd);
  const double gam = ksqrt(-1.0 / (m.gu00 + m.gu11 * vr * vr + m.gu33 * ell * ell - 2.0 * m.gu03 * ell));
  const double u_0 = -gam, u_1 = vr * gam, u_3 = ell * gam;
  return V4{{m.gu00 * u_0 + m.gu03 * u_3, m.gu11 * u_1, 0.0, m.gu33 * u_3 + m.gu03 * u_0}};
}

// u_from_u123, 'fromfile' (velocities.py:658-698), equatorial only
KG_INLINE V4 u_fromfile(double a, double r, const double* t, int64_t n) {
  const double u1 = table_interp(t, n, 1, r);
  const double u2 = table_interp(t, n, 2, r);
  const double u3 = table_interp(t, n, 3, r);
  const Angle q = equator();
  const double r2 = r * r, a2 = a * a;
  const double Delta = r2 - 2.0 * r + a2;
  const double Sigma = r2 + a2 * (q.cth * q.cth);
  const double g11 = Sigma / Delta, g22 = Sigma;
  const double as = a * q.sth;
  const double g33 = (r2 + a2 + 2.0 * r * (as * as) / Sigma) * (q.sth * q.sth);
  const double gcon00 = -(r2 + a2 + 2.0 * r * a2 / Sigma * (q.sth * q.sth)) / Delta;
  const double gcon03 = -2.0 * r * a / Sigma / Delta;
  const double alpha = 1.0 / ksqrt(-gcon00);
  const double gamma = ksqrt(1.0 + g11 * u1 * u1 + g22 * u2 * u2 + g33 * u3 * u3);
  return V4{{gamma / alpha, u1, u2, u3 - gamma * alpha * gcon03}};
}

// u_driftframe (velocities.py:487-623): drift velocity of the field
// f (numeric nu_parallel, no gammamax)
KG_INLINE V4 u_driftframe(double a, double r, const Angle& an, const kgeo_bfield_model& f, double nu) {
  const double a2 = a * a, r2 = r * r;
  const double cth2 = an.cth * an.cth, sth2 = an.sth * an.sth;
  const double Delta = r2 - 2.0 * r + a2;
  const double Sigma = r2 + a2 * cth2;
  const double Xi = (r2 + a2) * (r2 + a2) - Delta * a2 * sth2;
  const double omegaz = 2.0 * a * r / Xi;
  const double gdet = Sigma * an.sth;
  const double g11 = Sigma / Delta, g22 = Sigma, g33 = Xi * sth2 / Sigma;
  const double alpha = ksqrt(Delta * Sigma / Xi);
  const double eta3 = 2.0 * a * r / ksqrt(Delta * Sigma * Xi);
  const BLab b = bfield_lab(f, a, r, an);
  const double E1 = (b.omega - omegaz) * Xi * an.sth * b.B2 / Sigma;
  const double E2 = -(b.omega - omegaz) * Xi * an.sth * b.B1 / (Sigma * Delta);
  const double Bsq = g11 * b.B1 * b.B1 + g22 * b.B2 * b.B2 + g33 * b.B3 * b.B3;
  const double Esq = g11 * E1 * E1 + g22 * E2 * E2;
  const double B1c = g11 * b.B1, B2c = g22 * b.B2, B3c = g33 * b.B3;
  const double E1c = g11 * E1, E2c = g22 * E2, E3c = 0.0;
  const double k = alpha / (Bsq * gdet);
  const double vperp1 = k * (E2c * B3c - B2c * E3c);
  const double vperp2 = k * (E3c * B1c - B3c * E1c);
  const double vperp3 = k * (E1c * B2c - B1c * E2c);
  const double vpar_max = ksqrt(1.0 - Esq / Bsq);
  const double sB = ksqrt(Bsq);
  const double v1 = vperp1 + nu * vpar_max * b.B1 / sB;
  const double v2 = vperp2 + nu * vpar_max * b.B2 / sB;
  const double v3 = vperp3 + nu * vpar_max * b.B3 / sB;
  const double vsq = g11 * v1 * v1 + g22 * v2 * v2 + g33 * v3 * v3;
  const double gamma = 1.0 / ksqrt(1.0 - vsq);
  return V4{{gamma / alpha, gamma * v1, gamma * v2, gamma * (v3 + eta3)}};
}

// Velocity.u_lab (velocities.py:70-96)
KG_INLINE V4 velocity_u(const kgeo_velocity_model& v, double a, double r, const Angle& an) {
  const double s = v.retrograde ? -1.0 : 1.0;
  switch (v.kind) {
    case KGEO_VEL_ZAMO: return u_zamo(a, r);
    case KGEO_VEL_INFALL: return u_infall(a, r);
    case KGEO_VEL_KEP: return u_kepler(a, r, s);
    case KGEO_VEL_SUBKEP: return u_subkep(a, r, v.fac_subkep, s);
    case KGEO_VEL_GENERAL: return u_general(a, r, v.fac_subkep, v.beta_phi, v.beta_r, s);
    case KGEO_VEL_GELLES: return u_gelles(a, r, v.gelles_beta, v.gelles_chi);
    case KGEO_VEL_SIMFIT: return u_simfit(a, r, v.ell_isco, v.vr_isco, v.p1, v.p2, v.dd);
    case KGEO_VEL_FROMFILE: return u_fromfile(a, r, v.table, v.table_n);
    default: return u_driftframe(a, r, an, v.drift_field, v.nu_parallel);
  }
}

// Velocity.tetrades (velocities.py:124-176): e_(t), e_(x), e_(y), e_(z).
struct Tetrad {
  V4 e[4];
};

KG_INLINE Tetrad tetrad_of(const Metric& m, const V4& u) {
  const V4 ul = lower(m, u);
  const double uu03 = ul.v[0] * u.v[0] + ul.v[3] * u.v[3];
  const double Nr = ksqrt(-m.g11 * uu03 * (1.0 + ul.v[2] * u.v[2]));
  const double Nth = ksqrt(m.g22 * (1.0 + ul.v[2] * u.v[2]));
  const double Nph = ksqrt(-m.Delta * m.sth2 * uu03);
  Tetrad t;
  t.e[0] = V4{{-u.v[0], -u.v[1], -u.v[2], -u.v[3]}};
  t.e[1] = V4{{ul.v[1] * u.v[0] / Nr, -uu03 / Nr, 0.0, ul.v[1] * u.v[3] / Nr}};
  t.e[2] = V4{{ul.v[2] * u.v[0] / Nth, ul.v[2] * u.v[1] / Nth, (1.0 + ul.v[2] * u.v[2]) / Nth, ul.v[2] * u.v[3] / Nth}};
  t.e[3] = V4{{ul.v[3] / Nph, 0.0, 0.0, -ul.v[0] / Nph}};
  return t;
}

// Bfield.bfield_fluid (bfields.py:330-362): b^mu for ideal MHD in frame u.
KG_INLINE V4 fluid_field(const Metric& m, const BLab& B, const V4& u) {
  const V4 ul = lower(m, u);
  const double b0 = B.B1 * ul.v[1] + B.B2 * ul.v[2] + B.B3 * ul.v[3];
  return V4{{b0, (B.B1 + b0 * u.v[1]) / u.v[0], (B.B2 + b0 * u.v[2]) / u.v[0], (B.B3 + b0 * u.v[3]) / u.v[0]}};
}

// Bfield.bsq (bfields.py:365-395)
KG_INLINE double field_bsq(const kgeo_bfield_model& f, const kgeo_velocity_model& v, double a, double r,
                           const Angle& an) {
  const Metric m = metric_at(a, r, an);
  const V4 b = fluid_field(m, bfield_lab(f, a, r, an), velocity_u(v, a, r, an));
  return contract(b, lower(m, b));
}

// ---------------------------------------------------------------------------
// emissivities (emissivities.py)
// ---------------------------------------------------------------------------

// emissivities.py:141-157 (cgs constants :27-31)
KG_INLINE double j_nu_thermal(double ne, double B, double Te, double nu, double sin_thb) {
  constexpr double e = 4.80320425e-10, me = 9.10938356e-28, c = 2.99792458e10, kB = 1.380649e-16;
  const double Theta = kB * Te / (me * c * c);
  const double nu_B = e * B / (2.0 * kPiD * me * c);
  const double nu_c = 1.5 * nu_B * (Theta * Theta) * sin_thb;
  const double x = nu / kmax(nu_c, 1e-40);
  const double II = 2.5651 * (1.0 + 1.92 * kpow(x, -1.0 / 3.0) + 0.9977 * kpow(x, -2.0 / 3.0)) *
                    kexp(-1.8899 * kpow(x, 1.0 / 3.0));
  return (ne * (e * e) * nu) / (2.0 * ksqrt(3.0) * c * (Theta * Theta)) * II;
}

// Emissivity.jrest (emissivities.py:83-118); g, sinthb, bmag NaN-free
// optional arguments flagged by has_*.
KG_INLINE double jrest(const kgeo_emissivity_model& em, double a, double r, bool has_g, double g, bool has_s,
                       double sinthb, double bmag, double nu_obs, double specind) {
  if (em.kind == KGEO_EMIS_THERMAL) {
    const double nu_em = nu_obs / g;
    const double x = r / em.Rb;
    const double ne = em.ne0 * kpow(x, -em.alpha_n);
    const double Te = em.Te0 * kpow(x, -em.alpha_T);
    const double B = em.bfield_model ? bmag : em.B0 * kpow(x, -em.alpha_B);
    return j_nu_thermal(ne, B, Te, nu_em, sinthb);
  }
  double j;
  if (em.kind == KGEO_EMIS_CONSTANT) {
    j = 1.0;
  } else if (em.kind == KGEO_EMIS_BPL) {  // emisBPL :121-127
    const double rh = 1.0 + ksqrt(1.0 - a * a);
    const double l = klog(r / rh);
    j = kexp(em.p1 * l + em.p2 * (l * l));
  } else {  // emisGLM :130-138
    const double mu = (em.mu_ring != 0.0) ? em.mu_ring : 1.0 - ksqrt(1.0 - a * a);
    const double d = r - mu;
    const double z = em.gamma_off + kasinh(d / em.sigma);
    j = kexp(-0.5 * (z * z)) / ksqrt(d * d + em.sigma * em.sigma);
  }
  if (has_g) j = j * kpow(g, specind);
  if (has_s) j = j * kpow(sinthb, 1.0 + specind);
  return j;
}

// ---------------------------------------------------------------------------
// polarized observed intensity (equatorial_images.py:124-264)
// ---------------------------------------------------------------------------

struct Polarization {
  double sinthb, kappa_re, kappa_im, bsq;
};

// calc_polquantities (equatorial_images.py:336-487) for a lab-frame field:
// comoving-frame field and wave vector from the tetrad, the polarization
// vector f = k x B / |k|, and the Penrose-Walker constant kappa.
KG_INLINE Polarization pol_quantities(const kgeo_emission_model& M, double a, double r, double lam, double eta,
                                      double kr_sign, double kth_sign, const Angle& an) {
  const Metric m = metric_at(a, r, an);
  const V4 u = velocity_u(M.velocity, a, r, an);
  const Tetrad T = tetrad_of(m, u);
  const BLab B = bfield_lab(M.bfield, a, r, an);
  const V4 b = fluid_field(m, B, u);
  const V4 bl = lower(m, b);
  double Bp[3];
  for (int k = 0; k < 3; ++k) Bp[k] = contract(T.e[k + 1], bl);
  const double Bp_mag = ksqrt(Bp[0] * Bp[0] + Bp[1] * Bp[1] + Bp[2] * Bp[2]);
  const double lma = lam - a;
  const double R = (m.r2 + m.a2 - a * lam) * (m.r2 + m.a2 - a * lam) - m.Delta * (eta + lma * lma);
  const double TH = eta + m.a2 * m.cth2 - lam * lam * m.cth2 / m.sth2;
  const V4 kl{{-1.0, kr_sign * ksqrt(R) / m.Delta, kth_sign * ksqrt(TH), lam}};
  const double k0 = m.gu00 * kl.v[0] + m.gu03 * kl.v[3];
  const double k1 = m.gu11 * kl.v[1];
  const double k2 = m.gu22 * kl.v[2];
  const double k3 = m.gu33 * kl.v[3] + m.gu03 * kl.v[0];
  double kp[3];
  for (int k = 0; k < 3; ++k) kp[k] = contract(T.e[k + 1], kl);
  const double kp_mag = ksqrt(kp[0] * kp[0] + kp[1] * kp[1] + kp[2] * kp[2]);
  const double fx = (kp[1] * Bp[2] - kp[2] * Bp[1]) / kp_mag;
  const double fy = (kp[2] * Bp[0] - kp[0] * Bp[2]) / kp_mag;
  const double fz = (kp[0] * Bp[1] - kp[1] * Bp[0]) / kp_mag;
  Polarization P;
  P.sinthb = ksqrt(fx * fx + fy * fy + fz * fz) / Bp_mag;
  P.bsq = Bp_mag * Bp_mag;
  double f[4];
  for (int i = 0; i < 4; ++i) f[i] = T.e[1].v[i] * fx + T.e[2].v[i] * fy + T.e[3].v[i] * fz;
  const double A = k0 * f[1] - k1 * f[0] + a * m.sth2 * (k1 * f[3] - k3 * f[1]);
  const double Bk = ((m.r2 + m.a2) * (k3 * f[2] - k2 * f[3]) - a * (k0 * f[2] - k2 * f[0])) * an.sth;
  // kappa = (r - i a cos th) (A - i B)
  const double ac = a * an.cth;
  P.kappa_re = r * A - ac * Bk;
  P.kappa_im = -(r * Bk + ac * A);
  return P;
}

// Redshift with u^theta (equatorial_images.py:307-334)
KG_INLINE double redshift_u(const Metric& m, double a, double lam, double eta, double kr_sign, double kth_sign,
                            const V4& u) {
  const double lma = lam - a;
  const double R = (m.r2 + m.a2 - a * lam) * (m.r2 + m.a2 - a * lam) - m.Delta * (eta + lma * lma);
  const double TH = eta + m.a2 * m.cth2 - lam * lam * m.cth2 / m.sth2;
  return 1.0 / (1.0 * u.v[0] - lam * u.v[3] - np_sign(kth_sign) * u.v[2] * ksqrt(TH) -
                np_sign(kr_sign) * u.v[1] * ksqrt(R) / m.Delta);
}

struct StokesSub {
  double I, Q, U, g, sinthb, lp, Ie;
};

// One emitting subring of one pixel: Iobs after the masks
// (equatorial_images.py:186-257).  bsq0 is bfield.bsq at emissivity.Rb
// (the thermal normalisation, :226-232), unused otherwise.
KG_INLINE StokesSub stokes_subring(const kgeo_emission_model& M, double a, double sin_th_o, double alpha,
                                   double beta, double lam, double eta, double rs, double kr, double kth,
                                   double bsq0) {
  const Angle an = equator();
  const Metric m = metric_at(a, rs, an);
  const V4 u = velocity_u(M.velocity, a, rs, an);
  StokesSub o;
  o.g = redshift_u(m, a, lam, eta, kr, kth, u);
  double sinthb, bsq = 0.0, c2 = 0.0, s2 = 0.0;
  if (M.polarization) {
    const Polarization P = pol_quantities(M, a, rs, lam, eta, kr, kth, an);
    sinthb = P.sinthb;
    bsq = P.bsq;
    // calc_evpa (:524-544)
    const double mu = -(alpha + a * sin_th_o);
    const double k1 = P.kappa_re, k2 = P.kappa_im;
    const double p = beta * k1 + mu * k2;
    const double q = mu * k1 - beta * k2;
    const double den = (beta * beta + mu * mu) * (k1 * k1 + k2 * k2);
    c2 = (p * p - q * q) / den;
    s2 = (2.0 * p * q) / den;
  } else {
    sinthb = kPiD / 4.0;  // SINTHB (:20)
    if (M.emissivity.kind == KGEO_EMIS_THERMAL) {  // bfield.bsq (:212); only thermal reads it
      const V4 b = fluid_field(m, bfield_lab(M.bfield, a, rs, an), u);
      bsq = contract(b, lower(m, b));
    }
  }
  o.sinthb = sinthb;
  double bmag = ksqrt(bsq);
  if (M.emissivity.kind == KGEO_EMIS_THERMAL) bmag = (bmag / ksqrt(bsq0)) * M.emissivity.B0;
  double Iemis = jrest(M.emissivity, a, rs, true, o.g, true, sinthb, bmag, M.nu_obs, M.specind);
  if (M.pathlength && u.v[2] != 0.0) {
    // calc_pathlength_equatorial raises for u^theta != 0 (:509-510)
    o.I = o.Ie = o.lp = o.Q = o.U = NAN;
    return o;
  }
  if (M.pathlength) {
    // calc_pathlength_equatorial (:489-522)
    const double TH = eta + m.a2 * m.cth2 - lam * lam * m.cth2 / m.sth2;
    const double kyhat = -(kth * ksqrt(TH)) / rs;
    const double llp = (M.diskangle * rs) * ((1.0 / o.g) / kyhat);
    Iemis = Iemis * kfabs(llp);
    o.I = (o.g * o.g * o.g) * Iemis;
    o.Ie = Iemis;
    o.lp = llp;
  } else {
    o.I = (o.g * o.g) * Iemis;
    o.Ie = 0.0;
    o.lp = 0.0;
  }
  o.Q = M.polarization ? c2 * o.I : 0.0;
  o.U = M.polarization ? s2 * o.I : 0.0;
  return o;
}

// Whether the velocity can have u^theta != 0 in the equatorial plane, where
// calc_pathlength_equatorial raises (equatorial_images.py:509-510): the
// drift frame of a field with B^theta != 0 there.  (fromfile velocities
// carry their own U2 column, which the caller checks.)
KG_INLINE bool velocity_has_theta(const kgeo_velocity_model& v) {
  if (v.kind != KGEO_VEL_DRIFTFRAME) return false;
  switch (v.drift_field.kind) {
    case KGEO_BF_BZ_MONOPOLE:
    case KGEO_BF_BZ_PARA: return true;
    case KGEO_BF_POWER: return v.drift_field.p != 0.0;
    case KGEO_BF_GEN_POWER: return v.drift_field.p_val != 0.0;
    default: return false;
  }
}

// The models at one point (kgeo_model_eval): outputs at [k][n] + i for
// component k; NULL outputs and NULL g / sinthb / bmag (None) are skipped.
KG_INLINE void model_point(const kgeo_emission_model& M, double a, double r, const Angle& an, int64_t i, int64_t n,
                           const double* g, const double* sinthb, const double* bmag, double* u, double* tet,
                           double* B, double* omega, double* bsq, double* j) {
  if (u || tet || bsq) {
    const V4 uu = velocity_u(M.velocity, a, r, an);
    if (u)
      for (int k = 0; k < 4; ++k) u[k * n + i] = uu.v[k];
    const Metric m = metric_at(a, r, an);
    if (tet) {
      const Tetrad T = tetrad_of(m, uu);
      for (int row = 0; row < 4; ++row)
        for (int k = 0; k < 4; ++k) tet[(row * 4 + k) * n + i] = T.e[row].v[k];
    }
    if (bsq) {
      const V4 b = fluid_field(m, bfield_lab(M.bfield, a, r, an), uu);
      bsq[i] = contract(b, lower(m, b));
    }
  }
  if (B || omega) {
    const BLab b = bfield_lab(M.bfield, a, r, an);
    if (B) {
      B[i] = b.B1;
      B[n + i] = b.B2;
      B[2 * n + i] = b.B3;
    }
    if (omega) omega[i] = b.omega;
  }
  if (j)
    j[i] = jrest(M.emissivity, a, r, g != nullptr, g ? g[i] : 0.0, sinthb != nullptr, sinthb ? sinthb[i] : 0.0,
                 bmag ? bmag[i] : 0.0, M.nu_obs, M.specind);
}

struct StokesArrays {
  double *I, *Q, *U, *g, *r_s, *sin_thb, *Ir, *nmax_eq, *n_poloidal, *lp, *Ie;
};

// make_image's per-pixel loop over subrings (equatorial_images.py:76-122)
// for pixel i: Iobs of subrings 0..nsub-1 written at [i][m], N_max and
// nmax_poloidal at [i].  north: th_o < pi/2.  bsq0 as for stokes_subring.
KG_INLINE void stokes_pixel(const RayScalars<double>& c, bool north, const kgeo_emission_model& M, double alpha,
                            double beta, int nsub, double bsq0, int64_t i, const StokesArrays& o) {
  EqRay<double> e;
  eq_ray(c, alpha, beta, e);
  if (o.nmax_eq) o.nmax_eq[i] = e.vort ? -2.0 : e.N;
  if (o.n_poloidal) {
    CountConsts k;
    k.w = e.w;
    k.K = e.K;
    k.F0 = e.Fo;
    k.vort = e.vort;
    o.n_poloidal[i] = e.vort ? -2.0 : n_poloidal_at(k, e.Im);
  }
  const double rh = c.rplus;
  const kgeo_emissivity_model& em = M.emissivity;
  for (int m = 0; m < nsub; ++m) {
    const int64_t q = i * nsub + m;
    double Ir = -1.0, rs = -1.0;
    bool zero = true;
    if (!e.vort) {
      eq_subring(c, beta, e, m, Ir, rs);
      zero = (e.N < m) || (e.N == -1.0) || (rs <= rh) || (em.emiscut_in > 0.0 && rs < em.emiscut_in) ||
             (em.emiscut_out > 0.0 && rs > em.emiscut_out);
    }
    if (o.r_s) o.r_s[q] = rs;
    if (o.Ir) o.Ir[q] = Ir;
    StokesSub s{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!zero) {
      // radial_momentum_sign (:266-288), theta_momentum_sign (:290-305)
      const double kr = (!e.outside || Ir <= 0.5 * e.Im) ? 1.0 : -1.0;
      const double par = (m % 2 == 0) ? 1.0 : -1.0;
      const double kth = north ? -par : par;
      s = stokes_subring(M, c.a, c.sin_th, alpha, beta, e.lam, e.eta, rs, kr, kth, bsq0);
    }
    if (o.I) o.I[q] = s.I;
    if (o.Q) o.Q[q] = s.Q;
    if (o.U) o.U[q] = s.U;
    if (o.g) o.g[q] = s.g;
    if (o.sin_thb) o.sin_thb[q] = s.sinthb;
    if (o.lp) o.lp[q] = s.lp;
    if (o.Ie) o.Ie[q] = s.Ie;
  }
}

}  // namespace kg
