// FP64 special-function layer for the analytic Kerr raytracer.
//
// Replaces the reference's elliptic layer (kgeo/scipy_ellip_binding.py:87-237,
// kgeo/gsl_ellip_binding.py:154-237 and the direct scipy.special
// ellipkinc/ellipeinc/ellipj/ellipk calls listed there) with in-kernel code:
//
//   * Carlson symmetric integrals R_F, R_D, R_J, R_C by the duplication
//     theorem with the Taylor series of DLMF 19.36.1-2 (Carlson 1995,
//     Numer. Algorithms 10:13).
//   * Legendre F, E, Pi on a reduced amplitude (|phi| <= pi/2) given sin and
//     cos of the amplitude, plus the half-period count j so that
//     X(phi) = 2 j X_complete + X(phi - j pi)  (DLMF 19.2.10 periodicity).
//   * Jacobi sn, cn, dn by the descending Landen transformation
//     (DLMF 22.7.1-3) over the arithmetic-geometric-mean ladder
//     (DLMF 22.20(ii)): one sin/cos at the bottom, one division per level.
//
// Everything is templated on the scalar type T so the same source is used
// for the FP64 kernels (T = double) and for the forward-mode tangent kernels
// (T = kg::Dual<N>).  Control flow (iteration counts, branch choices) is
// decided on the value part only (kg::val()).

#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define KG_HD __host__ __device__
#define KG_INLINE __host__ __device__ inline __attribute__((always_inline))
#else
#define KG_HD
#define KG_INLINE inline
#endif

namespace kg {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kHalfPi = 1.570796326794896619231321691639751442;

// ---------------------------------------------------------------------------
// scalar wrappers for double (the Dual overloads live in kg_dual.hpp)
// ---------------------------------------------------------------------------
KG_INLINE double val(double x) { return x; }
KG_INLINE double ksqrt(double x) { return ::sqrt(x); }
KG_INLINE double kcbrt(double x) { return ::cbrt(x); }
KG_INLINE double ksin(double x) { return ::sin(x); }
KG_INLINE double kcos(double x) { return ::cos(x); }
KG_INLINE double ktan(double x) { return ::tan(x); }
KG_INLINE double kasin(double x) { return ::asin(x); }
KG_INLINE double kacos(double x) { return ::acos(x); }
KG_INLINE double katan(double x) { return ::atan(x); }
KG_INLINE double katan2(double y, double x) { return ::atan2(y, x); }
KG_INLINE double klog(double x) { return ::log(x); }
KG_INLINE double klog1p(double x) { return ::log1p(x); }
KG_INLINE double kexp(double x) { return ::exp(x); }
KG_INLINE double kfabs(double x) { return ::fabs(x); }
KG_INLINE double kpow(double x, double y) { return ::pow(x, y); }
KG_INLINE double kfloor(double x) { return ::floor(x); }
KG_INLINE double krint(double x) { return ::rint(x); }
KG_INLINE bool kfinite(double x) { return kfabs(x) < INFINITY; }  // false for NaN

// Waits until the wave's vector memory operations are complete (s_waitcnt
// vmcnt(0); nothing on the host).  A loop that reads registers loaded just
// before it calls this between the loads and the loop: the compiler's wait
// insertion otherwise keeps a wait for those loads inside the loop body, and
// as the counter completes in order, that wait then also holds each iteration
// until the previous iteration's stores have completed.
KG_INLINE void wait_vector_memory() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0) expcnt(7) lgkmcnt(15): vmcnt only
#endif
}

// True on every active lane of the wave when p holds on any of them (the
// host build evaluates one lane at a time: p itself).
KG_INLINE bool wave_any(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __any(p ? 1 : 0) != 0;
#else
  return p;
#endif
}

// Reciprocal to within an ulp or two, for the walk's well-scaled values:
// the hardware approximation refined by two Newton steps on the GPU (five
// instructions against the ten of a correctly rounded division), a plain
// division on the host.  Zero, infinite and NaN inputs give a non-finite
// result, which the callers' finiteness checks catch.
KG_INLINE double frcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
#else
  return 1.0 / x;
#endif
}
// arccos on [-1, 1] (NaN outside, as numpy's arccos), within 1 ulp, for
// the theta outputs: with it the theta walk holds four waves per SIMD (the
// libm call's registers kept it at three).  The classical half-angle reduction (e.g. Muller, Elementary
// Functions, ch. 11): for |x| < 1/2, acos x = pi/2 - asin x; otherwise
// acos |x| = 2 asin s, s = sqrt(z), z = (1 - |x|)/2 (exact), and
// acos(-|x|) = pi - acos |x|.  asin s = s + s z P(z) on z = s^2 in [0, 1/4],
// where P is a degree-11 Chebyshev fit of (asin sqrt z - sqrt z)/(z sqrt z)
// (mpmath, 50 digits; fit error 2.3e-16, so 6e-17 relative in asin),
// evaluated by Estrin's scheme.  Both ranges share one polynomial
// evaluation (selects, no branch).
template <class T> KG_INLINE T acos_poly(const T& x) {
  const bool small = kfabs(val(x)) < 0.5;  // false for NaN
  const T ax = kfabs(x);
  const T z = small ? x * x : 0.5 * (1.0 - ax);
  const T s = small ? ax : ksqrt(z);
  // Estrin: pairs, then z^2, z^4, z^8
  const T z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
  const T q0 = 0.1666666666666665 + z * 0.07500000000020764;
  const T q1 = 0.044642857103423646 + z * 0.03038194736709848;
  const T q2 = 0.02237204763174451 + z * 0.017355259955786323;
  const T q3 = 0.013929652902326633 + z * 0.011875494382636922;
  const T q4 = 0.0078029494773533175 + z * 0.01603551434914882;
  const T q5 = -0.010749050339697808 + z * 0.028169218060881414;
  const T p = (q0 + z2 * q1) + z4 * (q2 + z2 * q3) + z8 * (q4 + z2 * q5);
  const T r = s + s * (z * p);  // asin s
  // pi/2 and pi as a double plus its rounding error
  constexpr double kPio2Hi = 1.5707963267948966, kPio2Lo = 6.123233995736766e-17;
  const bool neg = val(x) < 0.0;
  if (small) return kPio2Hi - ((neg ? -r : r) - kPio2Lo);
  return neg ? 2.0 * kPio2Hi - (2.0 * r - 2.0 * kPio2Lo) : 2.0 * r;
}

KG_INLINE double kmax(double a, double b) { return a > b ? a : b; }
KG_INLINE double kmin(double a, double b) { return a < b ? a : b; }
KG_INLINE double parity_sign(double j) { return (::fmod(j, 2.0) == 0.0) ? 1.0 : -1.0; }

template <class T> KG_INLINE T sq(const T& x) { return x * x; }

// ---------------------------------------------------------------------------
// Carlson R_C(x, y), y != 0; Cauchy principal value for y < 0 (DLMF 19.2.20).
// Closed forms in terms of t = (y - x)/x:  R_C = x^{-1/2} g(t) with
// g(t) = atan(sqrt t)/sqrt t (t>0), atanh(sqrt(-t))/sqrt(-t) (-1<t<0)
// (DLMF 19.2.18-19), and its Maclaurin series near t = 0.
// ---------------------------------------------------------------------------
template <class T> KG_INLINE T rc_series(const T& t) {
  // g(t) = sum_{n>=0} (-t)^n / (2n+1), |t| < 2e-3: remainder < 1e-24
  return 1.0 + t * (-1.0 / 3.0 + t * (1.0 / 5.0 + t * (-1.0 / 7.0 + t * (1.0 / 9.0 +
         t * (-1.0 / 11.0 + t * (1.0 / 13.0 + t * (-1.0 / 15.0)))))));
}

// R_C(x, y) for x >= 0, y > 0 (DLMF 19.2.17-19).
template <class T> KG_INLINE T carlson_rc_pos(const T& x, const T& y) {
  if (val(x) == 0.0) return kHalfPi / ksqrt(y);
  T d = y - x;
  T t = d / x;
  if (kfabs(val(t)) < 2.0e-3) return rc_series(t) / ksqrt(x);
  if (val(d) > 0.0) {
    T sd = ksqrt(d);
    return katan(sd / ksqrt(x)) / sd;
  }
  // x > y: log((sqrt x + sqrt(x-y))/sqrt y)/sqrt(x-y), written with log1p:
  // (sqrt x - sqrt y) = (x-y)/(sqrt x + sqrt y)
  T sd = ksqrt(-d);
  T sx = ksqrt(x), sy = ksqrt(y);
  return klog1p((sd + (-d) / (sx + sy)) / sy) / sd;
}

// R_C(1, 1 + e), e > -1: the duplication terms of R_J.
template <class T> KG_INLINE T rc_1pe(const T& e) {
  const double ev = val(e);
  if (kfabs(ev) < 2.0e-3) return rc_series(e);
  if (ev > 0.0) {
    T s = ksqrt(e);
    return katan(s) / s;
  }
  // atanh(s)/s with s = sqrt(-e), atanh(s) = 0.5 log1p(2s/(1-s))
  T s = ksqrt(-e);
  return 0.5 * klog1p(2.0 * s / (1.0 - s)) / s;
}

// R_C(x, y), y != 0, with the Cauchy principal value for y < 0 (DLMF 19.2.20).
template <class T> KG_INLINE T carlson_rc(const T& x, const T& y) {
  if (val(y) > 0.0) return carlson_rc_pos(x, y);
  T xm = x - y;
  return ksqrt(x / xm) * carlson_rc_pos(xm, -y);
}

// ---------------------------------------------------------------------------
// Carlson R_F(x, y, z), x,y,z >= 0, at most one zero.
// Duplication (DLMF 19.26.18) until the normalised deviations are below
// ~2^-7, then the degree-7 series of DLMF 19.36.1.
// ---------------------------------------------------------------------------
template <class T> KG_INLINE T carlson_rf(T x, T y, T z) {
  T A = (x + y + z) * (1.0 / 3.0);
  const T dx0 = A - x, dy0 = A - y;
  // Stop when 4^-m * max|A0 - x_i| <= tol * |A_m|; the truncated series
  // error is then O(tol^8).
  double Q = kmax(kfabs(val(dx0)), kmax(kfabs(val(dy0)), kfabs(val(A - z)))) * 150.0;
  double f = 1.0;
  for (int it = 0; it < 16; ++it) {
    if (Q * f <= kfabs(val(A))) break;
    T sx = ksqrt(x), sy = ksqrt(y), sz = ksqrt(z);
    T lam = sx * (sy + sz) + sy * sz;
    x = (x + lam) * 0.25;
    y = (y + lam) * 0.25;
    z = (z + lam) * 0.25;
    A = (A + lam) * 0.25;
    f *= 0.25;
  }
  T X = dx0 * f / A;
  T Y = dy0 * f / A;
  T Z = -(X + Y);
  T E2 = X * Y - Z * Z;
  T E3 = X * Y * Z;
  T s = 1.0 + E2 * (-1.0 / 10.0 + E2 * (1.0 / 24.0 - E2 * (5.0 / 208.0))) +
        E3 * (1.0 / 14.0 + E3 * (3.0 / 104.0) + E2 * (-3.0 / 44.0 + E2 * (1.0 / 16.0)));
  return s / ksqrt(A);
}

// Degree-7 Taylor polynomial shared by R_J and R_D (DLMF 19.36.2 extended):
// E_k are elementary symmetric polynomials of the five normalised deviations
// (X, Y, Z, P, P) for R_J and (X, Y, Z, Z, Z) for R_D; the coefficients are
// 3/(3+2N) times the t^N coefficient of (1 + E2 t^2 - E3 t^3 + E4 t^4 - E5 t^5)^(-1/2).
template <class T> KG_INLINE T rj_series(const T& E2, const T& E3, const T& E4, const T& E5) {
  T E22 = E2 * E2;
  return 1.0 - (3.0 / 14.0) * E2 + (1.0 / 6.0) * E3 + (9.0 / 88.0) * E22 - (3.0 / 22.0) * E4 -
         (9.0 / 52.0) * E2 * E3 + (3.0 / 26.0) * E5 - (1.0 / 16.0) * E22 * E2 + (3.0 / 40.0) * E3 * E3 +
         (3.0 / 20.0) * E2 * E4 + (45.0 / 272.0) * E22 * E3 - (9.0 / 68.0) * (E3 * E4 + E2 * E5);
}

// ---------------------------------------------------------------------------
// Carlson R_J(x, y, z, p), p > 0 (DLMF 19.36.2 with the Carlson 1995
// summation of R_C(1, 1 + e_m) terms).  R_D(x,y,z) = R_J(x,y,z,z).
// ---------------------------------------------------------------------------
template <class T> KG_INLINE T carlson_rj_pos(T x, T y, T z, T p) {
  T A = (x + y + z + 2.0 * p) * 0.2;
  const T dx0 = A - x, dy0 = A - y, dz0 = A - z;
  const T delta = (p - x) * (p - y) * (p - z);
  double Q = kmax(kmax(kfabs(val(dx0)), kfabs(val(dy0))), kmax(kfabs(val(dz0)), kfabs(val(A - p)))) * 150.0;
  double f = 1.0;   // 4^-m
  double f3 = 1.0;  // 4^-3m
  T sum = 0.0;
  for (int it = 0; it < 16; ++it) {
    if (Q * f <= kfabs(val(A))) break;
    T sx = ksqrt(x), sy = ksqrt(y), sz = ksqrt(z), sp = ksqrt(p);
    T lam = sx * (sy + sz) + sy * sz;
    T d = (sp + sx) * (sp + sy) * (sp + sz);
    T e = (delta * f3) / (d * d);
    sum = sum + f * rc_1pe(e) / d;
    x = (x + lam) * 0.25;
    y = (y + lam) * 0.25;
    z = (z + lam) * 0.25;
    p = (p + lam) * 0.25;
    A = (A + lam) * 0.25;
    f *= 0.25;
    f3 *= (1.0 / 64.0);
  }
  T X = dx0 * f / A;
  T Y = dy0 * f / A;
  T Z = dz0 * f / A;
  T P = (X + Y + Z) * (-0.5);
  T P2 = P * P;
  T E2 = X * Y + X * Z + Y * Z - 3.0 * P2;
  T E3 = X * Y * Z + 2.0 * E2 * P + 4.0 * P * P2;
  T E4 = (2.0 * X * Y * Z + E2 * P + 3.0 * P * P2) * P;
  T E5 = X * Y * Z * P2;
  T s = rj_series(E2, E3, E4, E5);
  T sA = ksqrt(A);
  return f * s / (A * sA) + 6.0 * sum;
}

// R_D(x,y,z) = R_J(x,y,z,z): the R_C terms reduce to 1/(sqrt(z)(z+lam)).
template <class T> KG_INLINE T carlson_rd(T x, T y, T z) {
  T A = (x + y + 3.0 * z) * 0.2;
  const T dx0 = A - x, dy0 = A - y;
  double Q = kmax(kmax(kfabs(val(dx0)), kfabs(val(dy0))), kfabs(val(A - z))) * 150.0;
  double f = 1.0;
  T sum = 0.0;
  for (int it = 0; it < 16; ++it) {
    if (Q * f <= kfabs(val(A))) break;
    T sx = ksqrt(x), sy = ksqrt(y), sz = ksqrt(z);
    T lam = sx * (sy + sz) + sy * sz;
    sum = sum + f / (sz * (z + lam));
    x = (x + lam) * 0.25;
    y = (y + lam) * 0.25;
    z = (z + lam) * 0.25;
    A = (A + lam) * 0.25;
    f *= 0.25;
  }
  T X = dx0 * f / A;
  T Y = dy0 * f / A;
  T Z = (X + Y) * (-1.0 / 3.0);
  T XY = X * Y, Z2 = Z * Z;
  T E2 = XY - 6.0 * Z2;
  T E3 = (3.0 * XY - 8.0 * Z2) * Z;
  T E4 = 3.0 * (XY - Z2) * Z2;
  T E5 = XY * Z2 * Z;
  T s = rj_series(E2, E3, E4, E5);
  T sA = ksqrt(A);
  return f * s / (A * sA) + 3.0 * sum;
}

// R_J for any real rho != 0 with the reference's DLMF 19.20.14 transform
// for rho <= 0 (Cauchy principal value), as scipy_ellip_binding.py:207-230.
// rf must be R_F(x, y, z).
template <class T>
KG_INLINE T carlson_rj_any(const T& x, const T& y, const T& z, const T& rho, const T& rf) {
  if (val(rho) > 0.0) return carlson_rj_pos(x, y, z, rho);
  T q = -rho;
  T p = (z * (x + y + q) - x * y) / (z + q);
  T rj0 = carlson_rj_pos(x, y, z, p);
  T xy = x * y;
  T pq = p * q;
  T rc = carlson_rc(xy + pq, pq);
  T fac = 3.0 * ksqrt((xy * z) / (xy + pq));
  return ((p - z) * rj0 - 3.0 * rf + fac * rc) / (q + z);
}

// ---------------------------------------------------------------------------
// Legendre integrals on a reduced amplitude.
//
// An amplitude phi is carried as (s, c, j): phi = j*pi + phi', |phi'|<=pi/2,
// s = sin(phi'), c = cos(phi') >= 0.  Then (DLMF 19.2.10, 19.25.5-14)
//   F(phi|m)   = 2j K(m)      + s R_F(c^2, 1 - m s^2, 1)
//   E(phi|m)   = 2j E(m)      + s R_F - (m/3) s^3 R_D(c^2, 1 - m s^2, 1)
//   Pi(n;phi|m)= 2j Pi(n|m)   + s R_F + (n/3) s^3 R_J(c^2, 1 - m s^2, 1, 1 - n s^2)
// ---------------------------------------------------------------------------
template <class T> struct Amp {
  T s, c;    // sin, cos of the reduced amplitude
  double j;  // half-period count
};

template <class T> struct LegendreBase {
  T s, x, y, rf;  // s, cos^2, 1 - m s^2, R_F(x, y, 1)
  double j;
};

template <class T> KG_INLINE LegendreBase<T> legendre_base(const Amp<T>& A, const T& m) {
  LegendreBase<T> L;
  L.s = A.s;
  L.j = A.j;
  L.x = A.c * A.c;
  L.y = 1.0 - m * (A.s * A.s);
  L.rf = carlson_rf(L.x, L.y, T(1.0));
  return L;
}

// 2 j X_complete, which is 0 for j = 0 even where X_complete is infinite
template <class T> KG_INLINE T periods(double j, const T& Xc) { return (j != 0.0) ? (2.0 * j) * Xc : T(0.0); }

template <class T> KG_INLINE T legendre_F(const LegendreBase<T>& L, const T& Kc) {
  return periods(L.j, Kc) + L.s * L.rf;
}

// E - F without the R_F cancellation
template <class T>
KG_INLINE T legendre_EmF(const LegendreBase<T>& L, const T& m, const T& EmKc) {
  T rd = carlson_rd(L.x, L.y, T(1.0));
  return periods(L.j, EmKc) - (m * (1.0 / 3.0)) * (L.s * L.s * L.s) * rd;
}

template <class T>
KG_INLINE T legendre_Pi(const LegendreBase<T>& L, const T& n, const T& Pic) {
  T s2 = L.s * L.s;
  T rho = 1.0 - n * s2;
  T rj = carlson_rj_any(L.x, L.y, T(1.0), rho, L.rf);
  return periods(L.j, Pic) + L.s * L.rf + (n * (1.0 / 3.0)) * (s2 * L.s) * rj;
}

// ---------------------------------------------------------------------------
// Fused Carlson family at z = 1: R_F(x,y,1), R_D(x,y,1) and R_J(x,y,1,rho_i)
// for NP values rho_i (any sign; rho <= 0 via the reference's DLMF 19.20.14
// transform), sharing one duplication sequence of (x, y, z).  Per
// duplication step this costs 3 + NP square roots instead of 3 + 3 + 4 NP
// for separate evaluations.  The step count is the largest any member needs.
// ---------------------------------------------------------------------------
template <class T, int NP> struct CarlsonSet {
  T rf, rd;
  T rj[NP];
};

template <class T, int NP>
KG_INLINE void carlson_family(const T& x0, const T& y0, const T (&rho)[NP], bool want_rd, CarlsonSet<T, NP>& out) {
  const T one = T(1.0);
  T p0[NP], q[NP];
  bool neg[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    neg[i] = !(val(rho[i]) > 0.0);
    if (neg[i]) {
      q[i] = -rho[i];
      p0[i] = (x0 + y0 + q[i] - x0 * y0) / (one + q[i]);
    } else {
      q[i] = T(0.0);
      p0[i] = rho[i];
    }
  }
  // means and initial deviations of each member
  const T AF0 = (x0 + y0 + one) * (1.0 / 3.0);
  const T AD0 = (x0 + y0 + 3.0) * 0.2;
  T AJ0[NP], delta[NP];
  double Q = kmax(kmax(kfabs(val(AF0 - x0)), kfabs(val(AF0 - y0))), kfabs(val(AF0 - one))) / kfabs(val(AF0));
  if (want_rd) Q = kmax(Q, kmax(kmax(kfabs(val(AD0 - x0)), kfabs(val(AD0 - y0))), kfabs(val(AD0 - one))) / kfabs(val(AD0)));
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    AJ0[i] = (x0 + y0 + one + 2.0 * p0[i]) * 0.2;
    delta[i] = (p0[i] - x0) * (p0[i] - y0) * (p0[i] - one);
    const double a = val(AJ0[i]);
    Q = kmax(Q, kmax(kmax(kfabs(val(AJ0[i] - x0)), kfabs(val(AJ0[i] - y0))),
                     kmax(kfabs(val(AJ0[i] - one)), kfabs(val(AJ0[i] - p0[i])))) / kfabs(a));
  }
  // The means only shrink by the same lambda-driven map, so one relative
  // criterion on the initial spread bounds every member: stop when
  // 4^-m * Q <= 1/150 * (A_m / A_0), checked on the slowest-shrinking mean.
  Q *= 150.0;
  T x = x0, y = y0, z = one;
  T p[NP];
  T AF = AF0, AD = AD0, AJ[NP], sJ[NP];
  T sD = T(0.0);
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    p[i] = p0[i];
    AJ[i] = AJ0[i];
    sJ[i] = T(0.0);
  }
  double f = 1.0, f3 = 1.0;
  for (int it = 0; it < 16; ++it) {
    // member-wise criterion Q_rel * |A_0| * 4^-m <= |A_m|
    bool done = Q * f * kfabs(val(AF0)) <= kfabs(val(AF));
    if (want_rd) done = done && (Q * f * kfabs(val(AD0)) <= kfabs(val(AD)));
#pragma unroll
    for (int i = 0; i < NP; ++i) done = done && (Q * f * kfabs(val(AJ0[i])) <= kfabs(val(AJ[i])));
    if (done) break;
    T sx = ksqrt(x), sy = ksqrt(y), sz = ksqrt(z);
    T lam = sx * (sy + sz) + sy * sz;
    if (want_rd) sD = sD + f / (sz * (z + lam));
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      T sp = ksqrt(p[i]);
      T d = (sp + sx) * (sp + sy) * (sp + sz);
      T e = (delta[i] * f3) / (d * d);
      sJ[i] = sJ[i] + f * rc_1pe(e) / d;
      p[i] = (p[i] + lam) * 0.25;
      AJ[i] = (AJ[i] + lam) * 0.25;
    }
    x = (x + lam) * 0.25;
    y = (y + lam) * 0.25;
    z = (z + lam) * 0.25;
    AF = (AF + lam) * 0.25;
    AD = (AD + lam) * 0.25;
    f *= 0.25;
    f3 *= (1.0 / 64.0);
  }
  {  // R_F series (DLMF 19.36.1)
    T X = (AF0 - x0) * f / AF;
    T Y = (AF0 - y0) * f / AF;
    T Z = -(X + Y);
    T E2 = X * Y - Z * Z;
    T E3 = X * Y * Z;
    T sr = 1.0 + E2 * (-1.0 / 10.0 + E2 * (1.0 / 24.0 - E2 * (5.0 / 208.0))) +
           E3 * (1.0 / 14.0 + E3 * (3.0 / 104.0) + E2 * (-3.0 / 44.0 + E2 * (1.0 / 16.0)));
    out.rf = sr / ksqrt(AF);
  }
  if (want_rd) {
    T X = (AD0 - x0) * f / AD;
    T Y = (AD0 - y0) * f / AD;
    T Z = (X + Y) * (-1.0 / 3.0);
    T XY = X * Y, Z2 = Z * Z;
    T sr = rj_series(XY - 6.0 * Z2, (3.0 * XY - 8.0 * Z2) * Z, 3.0 * (XY - Z2) * Z2, XY * Z2 * Z);
    out.rd = f * sr / (AD * ksqrt(AD)) + 3.0 * sD;
  } else {
    out.rd = T(0.0);
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    T X = (AJ0[i] - x0) * f / AJ[i];
    T Y = (AJ0[i] - y0) * f / AJ[i];
    T Z = (AJ0[i] - one) * f / AJ[i];
    T P = (X + Y + Z) * (-0.5);
    T P2 = P * P;
    T E2 = X * Y + X * Z + Y * Z - 3.0 * P2;
    T XYZ = X * Y * Z;
    T sr = rj_series(E2, XYZ + 2.0 * E2 * P + 4.0 * P * P2, (2.0 * XYZ + E2 * P + 3.0 * P * P2) * P, XYZ * P2);
    T rjp = f * sr / (AJ[i] * ksqrt(AJ[i])) + 6.0 * sJ[i];
    if (neg[i]) {
      T xy = x0 * y0;
      T pq = p0[i] * q[i];
      T rc = carlson_rc(xy + pq, pq);
      T fac = 3.0 * ksqrt(xy / (xy + pq));
      out.rj[i] = ((p0[i] - one) * rjp - 3.0 * out.rf + fac * rc) / (q[i] + one);
    } else {
      out.rj[i] = rjp;
    }
  }
}

// Complete integrals K(m), E(m), Pi(n|m) (phi = pi/2: x = 0, y = 1 - m).
template <class T> KG_INLINE T ellip_K(const T& m) { return carlson_rf(T(0.0), 1.0 - m, T(1.0)); }

template <class T> KG_INLINE T ellip_EmK(const T& m) {
  return -(m * (1.0 / 3.0)) * carlson_rd(T(0.0), 1.0 - m, T(1.0));
}

template <class T> KG_INLINE T ellip_Pic(const T& n, const T& m, const T& Kc) {
  T rj = carlson_rj_any(T(0.0), 1.0 - m, T(1.0), 1.0 - n, Kc);
  return Kc + (n * (1.0 / 3.0)) * rj;
}

// Amplitude (s, c, j) from an angle phi (any real): j = rint(phi/pi).
template <class T> KG_INLINE Amp<T> amp_from_angle(const T& phi) {
  Amp<T> A;
  A.j = krint(val(phi) * (1.0 / kPi));
  T red = phi - A.j * kPi;
  A.s = ksin(red);
  A.c = kcos(red);
  return A;
}

// Amplitude of am(u|m) from sn, cn and the argument: am(u) lies in
// [j pi - pi/2, j pi + pi/2] for u in [(2j-1)K, (2j+1)K].
template <class T>
KG_INLINE Amp<T> amp_from_jacobi(const T& sn, const T& cn, double u, double Kc) {
  Amp<T> A;
  A.j = krint(u / (2.0 * Kc));
  const double sg = parity_sign(A.j);
  A.s = sg * sn;
  A.c = sg * cn;
  return A;
}

// ---------------------------------------------------------------------------
// Jacobi sn, cn, dn for 0 <= m < 1.
//
// AGM ladder a_0 = 1, b_0 = sqrt(1-m), a_{i+1} = (a_i+b_i)/2,
// b_{i+1} = sqrt(a_i b_i), c_{i+1} = (a_i - b_i)/2 and moduli k_i = c_i/a_i
// (descending Landen: k_{i+1} = (1-k_i')/(1+k_i')).  At the bottom level
// (k_N^2 below round-off) sn = sin(a_N u), cn = cos(a_N u), dn = 1, and the
// ascent uses DLMF 22.7.1-3 with w = a_i u:
//   den = 1 + k_i S^2,  sn = (1+k_i) S / den,  cn = C D / den,
//   dn = (1 - k_i S^2) / den.
// ---------------------------------------------------------------------------
constexpr int kMaxLadder = 12;
constexpr double kLadderTol = 7.450580596923828e-09;  // 2^-27: k_N^2 below round-off

template <class T> struct JacobiLadder {
  T k[kMaxLadder];  // k_1..k_N (index 0..N-1)
  T aN;             // a_N
  T K;              // K(m) = pi / (2 a_N)
  int n;            // N, or -1 for m >= 1 (sn = tanh, cn = dn = sech)
  int capped;       // ladder cap reached before convergence
};

template <class T> KG_INLINE void jacobi_ladder(const T& m, JacobiLadder<T>& L) {
  L.n = 0;
  L.capped = 0;
  if (!(val(m) < 1.0)) {
    L.n = -1;
    L.aN = T(0.0);
    L.K = T(INFINITY);
#pragma unroll
    for (int i = 0; i < kMaxLadder; ++i) L.k[i] = T(0.0);
    return;
  }
  T a = 1.0;
  T b = ksqrt(1.0 - m);
  T c = ksqrt(m);
#pragma unroll
  for (int i = 0; i < kMaxLadder; ++i) {
    if (val(c) > kLadderTol * val(a)) {
      T an = 0.5 * (a + b);
      T bn = ksqrt(a * b);
      T cn = 0.5 * (a - b);
      L.k[i] = cn / an;
      a = an;
      b = bn;
      c = cn;
      L.n = i + 1;
    } else {
      L.k[i] = T(0.0);
    }
  }
  if (val(c) > kLadderTol * val(a)) L.capped = 1;
  L.aN = a;
  L.K = kHalfPi / a;
}

template <class T>
KG_INLINE void jacobi_eval(const JacobiLadder<T>& L, const T& u, T& sn, T& cn, T& dn) {
  if (L.n < 0) {
    T e = kexp(-2.0 * kfabs(u));
    T th = (1.0 - e) / (1.0 + e);
    sn = (val(u) < 0.0) ? -th : th;
    cn = 2.0 * ksqrt(e) / (1.0 + e);
    dn = cn;
    return;
  }
  T w = L.aN * u;
  T S = ksin(w), C = kcos(w), D = T(1.0);
#pragma unroll
  for (int i = kMaxLadder - 1; i >= 0; --i) {
    if (i < L.n) {
      const T& k = L.k[i];
      T kS2 = k * (S * S);
      T inv = 1.0 / (1.0 + kS2);
      T Sn = (1.0 + k) * S * inv;
      T Cn = C * D * inv;
      D = (1.0 - kS2) * inv;
      S = Sn;
      C = Cn;
    }
  }
  sn = S;
  cn = C;
  dn = D;
}

}  // namespace kg
