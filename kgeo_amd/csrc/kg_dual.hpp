// Forward-mode dual numbers with N tangent directions, for the tangent
// (adjoint) kernels: instantiating the raytracer templates with Dual<N>
// propagates d/d(parameter) through every operation of the forward
// computation, including the Carlson duplication, the Landen ladders and
// the root finder.  Branches and iteration counts follow the value part
// (kg::val), so the tangent is the exact derivative of the computation the
// FP64 kernel performs.

#pragma once

#include "kg_math.hpp"

namespace kgd {

template <int N> struct Dual {
  double v;
  double d[N];
  KG_HD Dual() : v(0.0) {
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
  KG_HD Dual(double x) : v(x) {
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
  KG_HD static Dual seed(double x, int k) {
    Dual r(x);
    r.d[k] = 1.0;
    return r;
  }
};

// chain rule helper: f(x) with f'(x) = df.  A zero tangent stays exactly
// zero even where f' is infinite (sqrt at 0 in the complete integrals,
// where x = cos^2(pi/2) = 0 does not depend on the parameters).
template <int N> KG_INLINE Dual<N> chain(const Dual<N>& x, double fv, double df) {
  Dual<N> r(fv);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (x.d[i] == 0.0) ? 0.0 : df * x.d[i];
  return r;
}

template <int N> KG_INLINE Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> KG_INLINE Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> KG_INLINE Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> KG_INLINE Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  const double inv = 1.0 / b.v;
  const double q = a.v * inv;
  Dual<N> r(q);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * inv;
  return r;
}
template <int N> KG_INLINE Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> KG_INLINE Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> KG_INLINE Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> KG_INLINE Dual<N> operator*(double a, const Dual<N>& b) {
  Dual<N> r(a * b.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int N> KG_INLINE Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> KG_INLINE Dual<N> operator+(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v += b;
  return r;
}
template <int N> KG_INLINE Dual<N> operator-(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v -= b;
  return r;
}
template <int N> KG_INLINE Dual<N> operator*(const Dual<N>& a, double b) { return b * a; }
template <int N> KG_INLINE Dual<N> operator/(const Dual<N>& a, double b) { return (1.0 / b) * a; }

template <int N> KG_INLINE double val(const Dual<N>& x) { return x.v; }

template <int N> KG_INLINE Dual<N> ksqrt(const Dual<N>& x) {
  const double s = ::sqrt(x.v);
  return chain(x, s, 0.5 / s);
}
template <int N> KG_INLINE Dual<N> kcbrt(const Dual<N>& x) {
  const double c = ::cbrt(x.v);
  return chain(x, c, 1.0 / (3.0 * c * c));
}
template <int N> KG_INLINE Dual<N> ksin(const Dual<N>& x) { return chain(x, ::sin(x.v), ::cos(x.v)); }
template <int N> KG_INLINE Dual<N> kcos(const Dual<N>& x) { return chain(x, ::cos(x.v), -::sin(x.v)); }
template <int N> KG_INLINE Dual<N> ktan(const Dual<N>& x) {
  const double t = ::tan(x.v);
  return chain(x, t, 1.0 + t * t);
}
template <int N> KG_INLINE Dual<N> kasin(const Dual<N>& x) {
  return chain(x, ::asin(x.v), 1.0 / ::sqrt((1.0 - x.v) * (1.0 + x.v)));
}
template <int N> KG_INLINE Dual<N> kacos(const Dual<N>& x) {
  return chain(x, ::acos(x.v), -1.0 / ::sqrt((1.0 - x.v) * (1.0 + x.v)));
}
template <int N> KG_INLINE Dual<N> katan(const Dual<N>& x) { return chain(x, ::atan(x.v), 1.0 / (1.0 + x.v * x.v)); }
template <int N> KG_INLINE Dual<N> katan2(const Dual<N>& y, const Dual<N>& x) {
  const double r2 = x.v * x.v + y.v * y.v;
  Dual<N> r(::atan2(y.v, x.v));
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / r2;
  return r;
}
template <int N> KG_INLINE Dual<N> klog(const Dual<N>& x) { return chain(x, ::log(x.v), 1.0 / x.v); }
template <int N> KG_INLINE Dual<N> klog1p(const Dual<N>& x) { return chain(x, ::log1p(x.v), 1.0 / (1.0 + x.v)); }
template <int N> KG_INLINE Dual<N> kexp(const Dual<N>& x) {
  const double e = ::exp(x.v);
  return chain(x, e, e);
}
template <int N> KG_INLINE Dual<N> kfabs(const Dual<N>& x) { return (x.v < 0.0) ? -x : x; }
template <int N> KG_INLINE Dual<N> frcp(const Dual<N>& x) { return 1.0 / x; }
template <int N> KG_INLINE Dual<N> kpow(const Dual<N>& x, double y) {
  const double p = ::pow(x.v, y);
  return chain(x, p, y * ::pow(x.v, y - 1.0));
}
// x^y with both dual
template <int N> KG_INLINE Dual<N> kpow(const Dual<N>& x, const Dual<N>& y) {
  const double p = ::pow(x.v, y.v);
  Dual<N> r(p);
  const double dx = y.v * ::pow(x.v, y.v - 1.0);
  const double dy = (x.v > 0.0) ? p * ::log(x.v) : 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = dx * x.d[i] + dy * y.d[i];
  return r;
}

}  // namespace kgd
