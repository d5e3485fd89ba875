// Lensed curves of constant equatorial radius (equatorial_lensing.py:198-293):
// the screen radius rho at which the mbar-th equatorial crossing of the ray
// with screen polar angle varphi lands on a given radius, and the critical
// curve.  Host-compilable (KG_INLINE), so the CPU test build runs the same
// code as the kernels.

#pragma once

#include "kg_equatorial.hpp"

namespace kg {

constexpr double kLensInf = 1.0e50;  // equatorial_lensing.py:18

// r_equatorial2 (equatorial_lensing.py:198-216): r_s of subring mbar for the
// screen point rho (cos varphi, sin varphi), with the reference's mapping of
// rays that have no such crossing (r_s <= 0): kLensInf outside the critical
// curve, 1 inside.  c must describe an observer at infinity (r_o = inf, as
// the reference's R0).
KG_INLINE double r_equatorial2(const RayScalars<double>& c, double rho, double cphi, double sphi, int mbar) {
  const double alpha = cphi * rho;
  const double beta = sphi * rho;
  EqRay<double> e;
  eq_ray(c, alpha, beta, e);
  double r;
  bool outside;
  if (e.vort) {
    r = -1.0;  // equatorial_lensing.py:119-123
    // is_outside_crit takes eta < 0 as inside; eta == 0 goes by its roots
    outside = false;
    if (!(e.eta < 0.0)) {
      PixelSetup<double> px;
      px.status = 0;
      radial_roots(c, e.lam, e.eta, px);
      outside = (px.ri[3] == 0.0) && (px.rr[3] > c.rplus);
    }
  } else {
    double Ir;
    eq_subring(c, beta, e, mbar, Ir, r);
    outside = e.outside;
  }
  if (r <= 0.0 && outside) return kLensInf;
  if (r <= 0.0) return 1.0;
  return r;
}

// The bracket rho_of_req searches (equatorial_lensing.py:229-238).
KG_INLINE void rho_bracket(int mbar, double req, double& lo, double& hi) {
  if (mbar == 0) {
    lo = 0.0;
    hi = req + 5.0;
  } else if (mbar == 1) {
    lo = 0.0;
    hi = 100.0;
  } else {
    lo = 1.0;
    hi = 10.0;
  }
}

// Root of f in [xa, xb] by Brent's method (R. P. Brent, Algorithms for
// Minimization without Derivatives, 1973, ch. 4): inverse quadratic or
// secant steps inside the bracket, bisection when they do not shrink it fast
// enough.  Stops when the bracket's half-width is within
// (xtol + rtol |b|) / 2, the convention of scipy.optimize.brentq (xtol =
// 2e-12, rtol = 4 eps by default), or after maxiter evaluations.
// Returns the status: 0 converged, 1 no sign change over [xa, xb] (brentq
// raises there), 2 not converged; *root is NaN unless 0.
enum : int { BRENT_OK = 0, BRENT_NO_BRACKET = 1, BRENT_NOT_CONVERGED = 2 };

template <class F>
KG_INLINE int brent_root(const F& f, double xa, double xb, double xtol, double rtol, int maxiter, double* root) {
  double a = xa, b = xb;
  double fa = f(a), fb = f(b);
  *root = NAN;
  if (fa == 0.0) {
    *root = a;
    return BRENT_OK;
  }
  if (fb == 0.0) {
    *root = b;
    return BRENT_OK;
  }
  if (!((fa > 0.0 && fb < 0.0) || (fa < 0.0 && fb > 0.0))) return BRENT_NO_BRACKET;
  double c = a, fc = fa;   // the contrapoint: f(b) and f(c) have opposite signs
  double d = b - a, e = d; // the last step and the one before
  for (int it = 0; it < maxiter; ++it) {
    if ((fb > 0.0) == (fc > 0.0)) {  // keep the root between b and c
      c = a;
      fc = fa;
      d = b - a;
      e = d;
    }
    if (kfabs(fc) < kfabs(fb)) {  // b is the best estimate so far
      a = b;
      b = c;
      c = a;
      fa = fb;
      fb = fc;
      fc = fa;
    }
    const double tol = 0.5 * (xtol + rtol * kfabs(b));
    const double m = 0.5 * (c - b);
    if (kfabs(m) <= tol || fb == 0.0) {
      *root = b;
      return BRENT_OK;
    }
    if (kfabs(e) >= tol && kfabs(fa) > kfabs(fb)) {
      double p, q;
      const double s = fb / fa;
      if (a == c) {  // secant
        p = 2.0 * m * s;
        q = 1.0 - s;
      } else {  // inverse quadratic interpolation
        const double qa = fa / fc, rb = fb / fc;
        p = s * (2.0 * m * qa * (qa - rb) - (b - a) * (rb - 1.0));
        q = (qa - 1.0) * (rb - 1.0) * (s - 1.0);
      }
      if (p > 0.0) q = -q;
      else p = -p;
      if (2.0 * p < kmin(3.0 * m * q - kfabs(tol * q), kfabs(e * q))) {
        e = d;  // the interpolation is accepted
        d = p / q;
      } else {
        d = m;  // bisection
        e = m;
      }
    } else {
      d = m;
      e = m;
    }
    a = b;
    fa = fb;
    b += (kfabs(d) > tol) ? d : (m > 0.0 ? tol : -tol);
    fb = f(b);
  }
  return BRENT_NOT_CONVERGED;
}

// rho_of_req_single (equatorial_lensing.py:249-262) for one angle:
// brentq(r_equatorial2 - req) over the reference's bracket.
KG_INLINE int rho_of_req_at(const RayScalars<double>& c, double req, int mbar, double varphi, double* rho) {
  const double cphi = kcos(varphi), sphi = ksin(varphi);
  double lo, hi;
  rho_bracket(mbar, req, lo, hi);
  auto f = [&](double x) { return r_equatorial2(c, x, cphi, sphi, mbar) - req; };
  return brent_root(f, lo, hi, 2.0e-12, 4.0 * 2.220446049250313e-16, 100, rho);
}

// critical_curve (equatorial_lensing.py:264-293) at one photon-orbit radius
// r: the screen point (alpha, beta >= 0) of the bound photon orbit there
// (NaN where it is not visible from th_o).
KG_INLINE void critical_point(double a, double sin_th, double tan_th, double cos_th, double r, double* alpha,
                              double* beta) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double a2 = a * a;
  const double lam = -(r * r * r + a2 + r * a2 - 3.0 * (r * r)) / (a * r - a);
  const double eta = (r * r) * (3.0 * (r * r) + a2 - lam * lam) / (r * r - a2);
  *alpha = -lam / sin_th;
  const double ac = a * cos_th;
  const double lt = lam / tan_th;
  *beta = ksqrt(eta + ac * ac - lt * lt);
}

}  // namespace kg
