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
// (where the 8-term series of R_C(1, 1 - z^2) is exact to rounding) and
// every stepped value finite.  Otherwise the sample is evaluated directly
// and the walk continues from there.  The loops also restart from a direct
// evaluation every kAnchorEvery samples, which bounds the accumulated
// rounding to a few tens of ulps.

#pragma once

#include "kg_geodesic.hpp"

namespace kg {

constexpr double kStepZ2 = 0.005;
constexpr int kAnchorEvery = 16;

// R_C(1, 1 - z2) = sum_k z2^k / (2k + 1), |z2| <= kStepZ2 (0.005^8 / 17 < 1e-19)
KG_INLINE double rc_one_series(double z2) {
  double g = 1.0 / 15.0;
  g = g * z2 + 1.0 / 13.0;
  g = g * z2 + 1.0 / 11.0;
  g = g * z2 + 1.0 / 9.0;
  g = g * z2 + 1.0 / 7.0;
  g = g * z2 + 1.0 / 5.0;
  g = g * z2 + 1.0 / 3.0;
  return g * z2 + 1.0;
}

// Jacobi state (sn, cn, dn), (E - F)/m and NP third-kind integrals at one argument.
template <int NP> struct WalkState {
  double s, c, d;
  double D;
  double P[NP];
};

// Per-pixel constants of a walk: the state at the step, the parameter and,
// per integral, n and n (m - n)(1 - n).
template <int NP> struct WalkStep {
  WalkState<NP> b;
  double m;
  double n[NP], w2[NP];
};

template <int NP>
KG_INLINE void walk_init(WalkStep<NP>& w, double m, const double (&n)[NP]) {
  w.m = m;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    w.n[i] = n[i];
    w.w2[i] = n[i] * (m - n[i]) * (1.0 - n[i]);
  }
}

// One step a -> a + du.  Returns false (leaving `out` unspecified) when the
// step is not well conditioned; the caller then evaluates directly.
template <int NP, bool INTEGRALS>
KG_INLINE bool walk_step(const WalkState<NP>& a, const WalkStep<NP>& w, WalkState<NP>& out) {
  const WalkState<NP>& b = w.b;
  const double ss = a.s * b.s;
  const double inv = 1.0 / (1.0 - w.m * ss * ss);
  const double S = (a.s * (b.c * b.d) + a.c * b.s * a.d) * inv;
  const double C = (a.c * b.c - ss * a.d * b.d) * inv;
  const double Dn = (a.d * b.d - w.m * ss * a.c * b.c) * inv;
  out.s = S;
  out.c = C;
  out.d = Dn;
  bool ok = kfinite(S) && kfinite(C) && kfinite(Dn);
  if (!INTEGRALS) return ok;
  const double q = ss * S;
  out.D = a.D + b.D - q;
  const double cd = ss * C * Dn;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    // 1 - n sn^2 at both ends and p must share a sign: no pole of the
    // integrand (n > 1: principal value) lies inside the step
    const double r0 = 1.0 - w.n[i] * (a.s * a.s);
    const double r1 = 1.0 - w.n[i] * (S * S);
    const double p = r1 + w.n[i] * cd;
    const double ip = 1.0 / p;
    const double nq = w.n[i] * q * ip;
    const double z2 = w.w2[i] * (q * ip) * (q * ip);
    ok = ok && ((r0 > 0.0) == (r1 > 0.0)) && ((p > 0.0) == (r1 > 0.0)) && (kfabs(z2) <= kStepZ2);
    out.P[i] = a.P[i] + b.P[i] + nq * rc_one_series(z2);
    ok = ok && kfinite(out.P[i]);
  }
  return ok && kfinite(out.D);
}

}  // namespace kg
