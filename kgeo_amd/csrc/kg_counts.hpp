// Per-pixel count formulas of kerr_raytracing_utils.py:551-826: equatorial
// crossings, poloidal orbits, angular turning points and the radial turning
// point.  Host-compilable (KG_INLINE), so the CPU test build evaluates the
// same code as the kernels.  All arithmetic here rounds as numpy does (no FMA
// contraction): the integer counts are floor() of these values and must
// match the reference bit for bit away from exact integer boundaries.

#pragma once

#include "kg_geodesic.hpp"

namespace kg {

// Constants shared by the count formulas for one pixel.
//   eta > 0 (ordinary):  w = sqrt(-a^2 u_-), F0 = F(arcsin(cos th_o / sqrt(u_+)) | u_+/u_-),
//                        K = F(pi/2 | u_+/u_-)            (:590-595, :651-657, :730-738)
//   eta <= 0 (vortical, n_angular_turnings only, :757-767):
//                        w = sqrt(a^2 u_-), F0 = F(Nu | 1 - u_+/u_-),
//                        Nu = arcsin(sqrt((cos^2 th_o - u_-) / (u_+ - u_-))), K = F(pi/2 | 1 - u_+/u_-)
// eta is taken without raytrace_ana's spin-zero fix, as these functions do.
struct CountConsts {
  double w, F0, K;
  bool vort;  // eta <= 0
};

KG_INLINE CountConsts count_consts(const RayScalars<double>& c, double al, double be) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  CountConsts k;
  const double lam = -al * c.sin_th;
  const double eta = (al * al - c.a2) * c.cos2_th + be * be;
  k.vort = eta <= 0.0;
  double up, um;
  uplus_uminus(c, lam, eta, up, um);
  double ratio, a2um;
  if (c.minspin) {
    ratio = 0.0 * um;
    a2um = -(eta + lam * lam);
  } else {
    ratio = up / um;
    a2um = c.a2 * um;
  }
  if (!k.vort) {
    k.K = ellip_K(ratio);
    k.F0 = legendre_F(legendre_base(amp_from_sin(c.cos_th / ksqrt(up)), ratio), k.K);
    k.w = ksqrt(-a2um);
  } else {
    const double mv = 1.0 - ratio;
    k.K = ellip_K(mv);
    k.F0 = legendre_F(legendre_base(amp_from_sin(ksqrt((c.cos2_th - um) / (up - um))), mv), k.K);
    k.w = ksqrt(a2um);
  }
  return k;
}

// n_equatorial_crossings (:605-684): floor-valued, -2 for eta <= 0.
// north: th_o < pi/2 (the reference's test at :666).
KG_INLINE double n_equatorial_at(const CountConsts& k, bool north, double be, double tau) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  if (k.vort) return -2.0;
  const double x = tau * k.w;
  const double twoK = 2.0 * k.K;
  if (north) return (be <= 0.0) ? kfloor((x - k.F0) / twoK) : kfloor((x + k.F0) / twoK) - 1.0;
  return (be >= 0.0) ? kfloor((x + k.F0) / twoK) : kfloor((x - k.F0) / twoK) - 1.0;
}

// n_poloidal_orbits (:551-603): (sqrt(-a^2 u_-) tau) / (4 K), -2 for eta <= 0.
// numpy divides the complex product by the real 4K as (x + 0j) * (1 / 4K)
// (Smith's algorithm), which is what is written here.
KG_INLINE double n_poloidal_at(const CountConsts& k, double tau) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  if (k.vort) return -2.0;
  return (k.w * tau) * (1.0 / (4.0 * k.K));
}

// n_angular_turnings (:686-792): (tau w -+ F0) / (2K) + 1/2 (minus for
// beta <= 0); the vortical beta <= 0 branch adds 1/2 more to values above
// 1/2, as the reference does (:778-781).
KG_INLINE double n_angular_at(const CountConsts& k, double be, double tau) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double x = tau * k.w;
  const double twoK = 2.0 * k.K;
  if (be <= 0.0) {
    double n = (x - k.F0) / twoK + 0.5;
    if (k.vort && n > 0.5) n += 0.5;
    return n;
  }
  return (x + k.F0) / twoK + 0.5;
}

// Radial turning point of one pixel: the ray lies outside the critical
// curve (is_outside_crit, :794-826: r_4 real and beyond r_+, eta >= 0) and
// then turns at half its total Mino time (mino_total, :518-546).  tau_tot
// is raytrace_ana's (with the spin-zero eta fix, kerr_raytracing_ana.py:80-88).
struct RadialTurn {
  bool outside;
  double half;  // 0.5 * tau_tot
};

KG_INLINE RadialTurn radial_turn_consts(const RayScalars<double>& c, double al, double be) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  RadialTurn rt;
  PixelSetup<double> raw;
  raw.status = 0;
  const double lam = -al * c.sin_th;
  const double eta = (al * al - c.a2) * c.cos2_th + be * be;
  radial_roots(c, lam, eta, raw);
  rt.outside = (raw.ri[3] == 0.0) && (raw.rr[3] > c.rplus) && !(eta < 0.0);
  PixelSetup<double> px;
  pixel_setup(c, al, be, px);
  rt.half = 0.5 * px.tau_tot;
  return rt;
}

// Radial turning points passed by Mino time tau: 1 past the turning point
// (not tau <= tau_tot / 2, the reference's own test, equatorial_images.py:284).
KG_INLINE int radial_turns_at(const RadialTurn& rt, double tau) {
  return (rt.outside && !(tau <= rt.half)) ? 1 : 0;
}

}  // namespace kg
