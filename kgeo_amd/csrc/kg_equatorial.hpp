// Equatorial-emission image: per pixel, the emission radius of the mbar-th
// equatorial crossing, its redshift for a Keplerian disk (Cunningham plunge
// inside the ISCO) and the observed intensity g^3 r_s^-p, summed over
// subrings mbar = 0..mbar_max.
//
// Behaviour follows the reference:
//   r_equatorial            equatorial_lensing.py:85-196 (F0, K, N_max, I_r, r_s)
//   zero mask               equatorial_images.py:175
//   radial_momentum_sign    equatorial_images.py:266-288 (is_outside_crit,
//                           kerr_raytracing_utils.py:794-826)
//   theta_momentum_sign     equatorial_images.py:290-305
//   u_kep                   velocities.py:229-281 (ISCO :242-244)
//   calc_redshift           equatorial_images.py:307-334
// The intensity g^3 r_s^-p is this workload's composition of those pieces
// (the reference's Iobs uses g^2 j or g^3 |l_p| j).
//
// Templated on the scalar type: T = double for the image kernel, T =
// kgd::Dual<3> for the tangent kernel (d/d a, d/d th_o, d/d p).

#pragma once

#include "kg_geodesic.hpp"

namespace kg {

constexpr int kMaxSubrings = 4;

template <class T> struct EqPixel {
  T image;
  T I[kMaxSubrings], g[kMaxSubrings], r_s[kMaxSubrings], Ir[kMaxSubrings];
  T Imax;
  int Nmax;        // -2 for eta <= 0 (the reference's vortical sentinel)
  int zero_mask;   // bit mbar set where subring mbar carries no emission
};

// ISCO radius, prograde (velocities.py:242-244)
template <class T> KG_INLINE T r_isco(const T& a) {
  T a2 = a * a;
  T z1 = 1.0 + kcbrt(1.0 - a2) * (kcbrt(1.0 + a) + kcbrt(1.0 - a));
  T z2 = ksqrt(3.0 * a2 + z1 * z1);
  return 3.0 + z2 - ksqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

// Keplerian 4-velocity outside the ISCO, Cunningham plunge inside
// (velocities.py:229-281, retrograde = False): returns u^t, u^r, u^phi.
template <class T>
KG_INLINE void u_kep(const T& a, const T& ri, const T& r, T& u0, T& u1, T& u3) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double asign = (val(a) > 0.0) ? 1.0 : ((val(a) < 0.0) ? -1.0 : 0.0);  // np.sign(a)
  const T spin = kfabs(a);
  if (val(r) >= val(ri)) {
    T r15 = kpow(r, 1.5);
    T Om = asign / (r15 + spin);
    u0 = (r15 + spin) / ksqrt(kpow(r, 3.0) - 3.0 * (r * r) + 2.0 * spin * r15);
    u1 = T(0.0);
    u3 = Om * u0;
  } else {
    T Delta = r * r - 2.0 * r + a * a;
    T sri = ksqrt(ri);
    T ell = asign * (ri * ri + a * a - 2.0 * spin * sri) / (kpow(ri, 1.5) - 2.0 * sri + spin);
    T gam = ksqrt(1.0 - 2.0 / (3.0 * ri));
    T H = (2.0 * r - a * ell) / Delta;
    u0 = gam * (1.0 + (2.0 / r) * (1.0 + H));
    u1 = -ksqrt(2.0 / (3.0 * ri)) * kpow(ri / r - 1.0, 1.5);
    u3 = gam * (ell + a * H) / (r * r);
  }
}

// cos(pi/2)^2 and sin(pi/2)^2 as numpy evaluates them (th = np.pi/2)
constexpr double kCos2HalfPi = 6.123233995736766e-17 * 6.123233995736766e-17;

// g = 1/(u^t - lam u^phi - sign(k_th) u^th sqrt(Theta) - sign(k_r) u^r sqrt(R)/Delta)
// at theta = pi/2 (equatorial_images.py:307-334), u^th = 0.
template <class T>
KG_INLINE T redshift(const T& a, const T& r, const T& lam, const T& eta, double kr_sign, const T& u0, const T& u1,
                     const T& u3) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  T a2 = a * a;
  T r2 = r * r;
  T Delta = r2 - 2.0 * r + a2;
  T lma = lam - a;
  T R = sq(r2 + a2 - a * lam) - Delta * (eta + lma * lma);
  T TH = eta + a2 * kCos2HalfPi - (lam * lam) * kCos2HalfPi;
  T th_term = 0.0 * ksqrt(TH);  // u^theta = 0; NaN for Theta < 0 as in numpy
  return 1.0 / (u0 - lam * u3 - th_term - (kr_sign * u1) * ksqrt(R) / Delta);
}

// The subring-independent part of r_equatorial for one pixel
// (equatorial_lensing.py:110-157): lambda, eta, K, F0, the roots, I_max,
// N_max and the radial constants; and is_outside_crit
// (kerr_raytracing_utils.py:821-823).  eta <= 0 (vortical) leaves the rest
// unset (equatorial_lensing.py:119-123).
template <class T> struct EqRay {
  T lam, eta, K, Fo, w, Im;
  double N;
  bool vort, outside;
  RadialConsts<T> rc;
  JacobiLadder<T> lad;
};

template <class T>
KG_INLINE void eq_ray(const RayScalars<T>& c, const T& alpha, const T& beta, EqRay<T>& e) {
  PixelSetup<T> px;
  px.status = 0;
  {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    e.lam = -alpha * c.sin_th;
    e.eta = (alpha * alpha - c.a2) * c.cos2_th + beta * beta;
  }
  const T& lam = e.lam;
  const T& eta = e.eta;
  e.vort = val(eta) <= 0.0;
  e.outside = false;
  e.N = -2.0;
  if (e.vort) return;
  T up, um, ratio, a2um;
  {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    uplus_uminus(c, lam, eta, up, um);
    if (c.minspin) {
      ratio = 0.0 * um;
      a2um = -(eta + lam * lam);
    } else {
      ratio = up / um;
      a2um = c.a2 * um;
    }
    e.K = ellip_K(ratio);
    e.Fo = legendre_F(legendre_base(amp_from_sin(c.cos_th / ksqrt(up)), ratio), e.K);
    radial_roots(c, lam, eta, px);
    e.Im = mino_total(c, px);
    e.w = ksqrt(-a2um);
    const double x = val(e.Im) * val(e.w);
    const double Fo = val(e.Fo), K = val(e.K);
    if (c.h > 0.0) {
      e.N = (val(beta) <= 0.0) ? kfloor((x - Fo) / (2.0 * K)) : kfloor((x + Fo) / (2.0 * K)) - 1.0;
    } else {
      e.N = (val(beta) >= 0.0) ? kfloor((x + Fo) / (2.0 * K)) : kfloor((x - Fo) / (2.0 * K)) - 1.0;
    }
  }
  radial_consts(c, px, false, e.rc);
  jacobi_ladder(e.rc.v[R_k], e.lad);
  e.outside = (val(px.ri[3]) == 0.0) && (val(px.rr[3]) > val(c.rplus)) && !(val(eta) < 0.0);
}

// I_r and r_s of subring m (equatorial_lensing.py:158-188): r_s = 0 where
// N_max < m.  Needs a non-vortical ray.
template <class T>
KG_INLINE void eq_subring(const RayScalars<T>& c, const T& beta, const EqRay<T>& e, int m, T& Ir, T& rs) {
  {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    double msub;
    if (c.h > 0.0) {
      msub = m + ((val(beta) > 0.0) ? 1.0 : 0.0);
      Ir = (val(beta) <= 0.0) ? (2.0 * msub * e.K + e.Fo) / e.w : (2.0 * msub * e.K - e.Fo) / e.w;
    } else {
      msub = m + ((val(beta) < 0.0) ? 1.0 : 0.0);
      Ir = (val(beta) >= 0.0) ? (2.0 * msub * e.K - e.Fo) / e.w : (2.0 * msub * e.K + e.Fo) / e.w;
    }
  }
  rs = radial_eval(e.rc, e.lad, Ir, false).r;
  if (e.N < m) rs = T(0.0);
}

// One pixel of the equatorial image for subrings 0..mbar_max.
template <class T>
KG_INLINE void equatorial_pixel(const RayScalars<T>& c, const T& alpha, const T& beta, const T& p_emis,
                                const T& ri, int mbar_max, EqPixel<T>& o) {
  o.image = T(0.0);
  o.zero_mask = 0;
  for (int m = 0; m < kMaxSubrings; ++m) {
    o.I[m] = T(0.0);
    o.g[m] = T(0.0);
    o.r_s[m] = T(0.0);
    o.Ir[m] = T(0.0);
  }
  EqRay<T> e;
  eq_ray(c, alpha, beta, e);
  if (e.vort) {  // vortical sentinels (equatorial_lensing.py:119-123)
    for (int m = 0; m <= mbar_max; ++m) {
      o.r_s[m] = T(-1.0);
      o.Ir[m] = T(-1.0);
      o.zero_mask |= 1 << m;
    }
    o.Imax = T(-1.0);
    o.Nmax = -2;
    return;
  }
  o.Imax = e.Im;
  o.Nmax = static_cast<int>(e.N);
  const T rh = c.rplus;
  for (int m = 0; m <= mbar_max; ++m) {
    T Ir, rs;
    eq_subring(c, beta, e, m, Ir, rs);
    o.Ir[m] = Ir;
    o.r_s[m] = rs;
    // equatorial_images.py:175 (a NaN r_s is not masked there)
    const bool zero = (e.N < m) || (e.N == -1.0) || (val(rs) <= val(rh));
    if (zero) {
      o.zero_mask |= 1 << m;
      continue;
    }
    const double kr = (!e.outside || val(Ir) <= 0.5 * val(e.Im)) ? 1.0 : -1.0;
    T u0, u1, u3;
    u_kep(c.a, ri, rs, u0, u1, u3);
    T g = redshift(c.a, rs, e.lam, e.eta, kr, u0, u1, u3);
    T I = (g * g * g) * kpow(rs, -p_emis);
    o.g[m] = g;
    o.I[m] = I;
    o.image = o.image + I;
  }
}

}  // namespace kg
