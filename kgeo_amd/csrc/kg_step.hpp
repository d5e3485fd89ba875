// Advancing the theta and radial states along the uniform Mino-time grid.
//
// On the grid tau_j = j h the Jacobi arguments are linear in j
// (u_j = u_0 + j du), so the state at u_{j+1} follows from the state at u_j
// and the state at the step du by the addition theorems instead of a fresh
// Landen ladder and Carlson duplication per sample:
//
//   sn, cn, dn (u + v)          DLMF 22.8.1-3
//   (E - F)/m:  D(u+v) = D(u) + D(v) - sn u sn v sn(u+v)          (DLMF 22.16.27
//               with E(u) = u + m D(u))
//   Pi(n; am u | m) =: P(u):
//     P(u+v) = P(u) + P(v) + (n q / p) R_C(1, 1 - z^2),
//     q = sn u sn v sn(u+v),  p = 1 - n sn^2(u+v) + n sn u sn v cn(u+v) dn(u+v),
//     z^2 = n (m - n)(1 - n) q^2 / p^2.
//
// The Pi relation is Jacobi's addition theorem for the third kind
// (Whittaker & Watson 22.74) rewritten for a real characteristic n =
// m sn^2(a): with sn a, cn a, dn a eliminated, the log of the theta-function
// ratio becomes artanh(z)/z = R_C(1, 1 - z^2) (arctan for z^2 < 0).  It is
// checked against mpmath quadrature in the test suite.
//
// A step is taken only when it is well conditioned: no pole of the integrand
// (1 - n sn^2 = 0, n > 1) between u and u + v or at p = 0, |z^2| <= kStepZ2
// (where the 15-term series of R_C(1, 1 - z^2) is exact to rounding) and
// every stepped value finite.  Otherwise the sample is evaluated directly
// and the walk continues from there.  The loops also restart from a direct
// evaluation every kAnchorEvery samples, which bounds the accumulated
// rounding to a few tens of ulps.

#pragma once

#include "kg_geodesic.hpp"

namespace kg {

constexpr double kStepZ2 = 0.04;
constexpr int kAnchorEvery = 16;

// R_C(1, 1 - z2) = sum_k z2^k / (2k + 1) for |z2| <= kStepZ2
// (0.04^15 / 31 < 1e-22: the truncation is below rounding)
template <class T> KG_INLINE T rc_one_series(const T& z2) {
  T g = T(1.0 / 29.0);
  g = g * z2 + 1.0 / 27.0;
  g = g * z2 + 1.0 / 25.0;
  g = g * z2 + 1.0 / 23.0;
  g = g * z2 + 1.0 / 21.0;
  g = g * z2 + 1.0 / 19.0;
  g = g * z2 + 1.0 / 17.0;
  g = g * z2 + 1.0 / 15.0;
  g = g * z2 + 1.0 / 13.0;
  g = g * z2 + 1.0 / 11.0;
  g = g * z2 + 1.0 / 9.0;
  g = g * z2 + 1.0 / 7.0;
  g = g * z2 + 1.0 / 5.0;
  g = g * z2 + 1.0 / 3.0;
  return g * z2 + 1.0;
}

// Jacobi state (sn, cn, dn), (E - F)/m and NP third-kind integrals at one
// argument.  T = double in the kernels; the op-counting CPU build
// instantiates it on its counting type.
template <class T, int NP> struct WalkState {
  T s, c, d;
  T D;
  T P[NP];
};

// Per-pixel constants of a walk: the state at the step, the parameter and,
// per integral, n and n (m - n)(1 - n).
template <class T, int NP> struct WalkStep {
  WalkState<T, NP> b;
  T m;
  T n[NP], w2[NP];  // the walked characteristic and n (m - n)(1 - n)
  T pt[NP];         // sqrt((n - 1)(1 - m/n)) of a transformed integral, else 0
  bool tr[NP];
};

// Characteristics n > 1 (principal value) are walked through
//   Pi(n; am u | m) = -Pi(m/n; am u | m) + u + L(u),
//   L(u) = ln|(1 + x)/(1 - x)| / (2 pt),  x = pt sn u / (cn u dn u),
//   pt = sqrt((n - 1)(1 - m/n))                           (DLMF 19.7.9)
// whose right-hand integral has 0 < m/n < 1 and no pole: the log carries the
// singularity and is evaluated per sample.  Used for 0 <= m < 1, n > 1 and
// pt >= kTransformMinPt (the log term amplifies rounding by ~1/pt); anywhere
// else n is walked as it is.
constexpr double kTransformMinPt = 0.05;

template <class T> KG_INLINE T pv_log_term(const T& pt, const T& sn, const T& cn, const T& dn) {
  const T x = pt * sn / (cn * dn);
  const T l = (kfabs(val(x)) < 1.0) ? klog1p(2.0 * x / (1.0 - x)) : klog1p(2.0 / (x - 1.0));
  return l / (2.0 * pt);
}

template <class T, int NP>
KG_INLINE void walk_init(WalkStep<T, NP>& w, const T& m, const T (&n)[NP]) {
  w.m = m;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const bool cand = (val(n[i]) > 1.0) && kfinite(val(n[i])) && (val(m) >= 0.0) && (val(m) < 1.0);
    const T pt = cand ? ksqrt((n[i] - 1.0) * (1.0 - m / n[i])) : T(0.0);
    w.tr[i] = cand && (val(pt) >= kTransformMinPt);
    const T nw = w.tr[i] ? m / n[i] : n[i];
    w.pt[i] = w.tr[i] ? pt : T(0.0);
    w.n[i] = nw;
    w.w2[i] = nw * (m - nw) * (1.0 - nw);
  }
}

// Pi(n; am u) -> the walked integral at the same argument, and back.
template <class T, int NP>
KG_INLINE T walked_pi(const WalkStep<T, NP>& w, int i, const T& Pn, const T& u, const T& sn, const T& cn,
                      const T& dn) {
  return w.tr[i] ? -Pn + u + pv_log_term(w.pt[i], sn, cn, dn) : Pn;
}

template <class T, int NP>
KG_INLINE T original_pi(const WalkStep<T, NP>& w, int i, const T& Pw, const T& u, const T& sn, const T& cn,
                        const T& dn) {
  return w.tr[i] ? -Pw + u + pv_log_term(w.pt[i], sn, cn, dn) : Pw;
}

// One step a -> a + du.  Returns false (leaving `out` unspecified) when the
// step is not well conditioned; the caller then evaluates directly.
template <class T, int NP, bool INTEGRALS>
KG_INLINE bool walk_step(const WalkState<T, NP>& a, const WalkStep<T, NP>& w, WalkState<T, NP>& out) {
  const WalkState<T, NP>& b = w.b;
  const T ss = a.s * b.s;
  const T inv = 1.0 / (1.0 - w.m * ss * ss);
  const T S = (a.s * (b.c * b.d) + a.c * b.s * a.d) * inv;
  const T C = (a.c * b.c - ss * a.d * b.d) * inv;
  const T Dn = (a.d * b.d - w.m * ss * a.c * b.c) * inv;
  out.s = S;
  out.c = C;
  out.d = Dn;
  bool ok = kfinite(val(S)) && kfinite(val(C)) && kfinite(val(Dn));
  if (!INTEGRALS) return ok;
  const T q = ss * S;
  out.D = a.D + b.D - q;
  const T cd = ss * C * Dn;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    // 1 - n sn^2 at both ends and p must share a sign: no pole of the
    // integrand (n > 1: principal value) lies inside the step
    const T r0 = 1.0 - w.n[i] * (a.s * a.s);
    const T r1 = 1.0 - w.n[i] * (S * S);
    const T p = r1 + w.n[i] * cd;
    const T ip = 1.0 / p;
    const T nq = w.n[i] * q * ip;
    const T z2 = w.w2[i] * (q * ip) * (q * ip);
    ok = ok && ((val(r0) > 0.0) == (val(r1) > 0.0)) && ((val(p) > 0.0) == (val(r1) > 0.0)) &&
         (kfabs(val(z2)) <= kStepZ2);
    out.P[i] = a.P[i] + b.P[i] + nq * rc_one_series(z2);
    ok = ok && kfinite(val(out.P[i]));
  }
  return ok && kfinite(val(out.D));
}

}  // namespace kg
