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
// (where at most 12 terms of the series of R_C(1, 1 - z^2) are exact to
// rounding) and
// every stepped value finite.  Otherwise the sample is evaluated directly
// and the walk continues from there.  The loops also restart from a direct
// evaluation every kAnchorEveryLong samples (theta, radial case 2) or
// kAnchorEvery samples (radial cases 3 and 4, whose R1/S1 blocks amplify
// rounding), which bounds the accumulated rounding to a few tens of ulps.

#pragma once

#include "kg_geodesic.hpp"

namespace kg {

constexpr double kStepZ2 = 0.04;
constexpr double kPoleGuard = 1.0e-4;
// |ln|(1 + x)/(1 - x)|| above kLogGuard (|x| within ~1e-4 of 1) means a
// walked log term sits next to its singularity, where the terms that cancel
// in the outputs are each amplified by ~e^|ln| / 2: such steps go direct.
constexpr double kLogGuard = 10.0;
constexpr int kAnchorEvery = 16;       // radial cases 3 and 4
constexpr int kAnchorEveryLong = 32;   // theta and radial case 2

// R_C(1, 1 - z2) = sum_k z2^k / (2k + 1), the first NT terms (NT even), by
// Estrin's scheme: pairs 1/(4k+1) + z2/(4k+3), then combined with z2^2,
// z2^4, ...  Its dependency depth is 1 + log2(NT) multiply-adds instead of
// Horner's NT - 1, which matters in the walk loops, where the series sits on
// each step's critical path.
template <int NT, class T> KG_INLINE T rc_one_terms(const T& z2) {
  static_assert(NT % 2 == 0, "even term count");
  T p[NT / 2];
#pragma unroll
  for (int k = 0; k < NT / 2; ++k) p[k] = 1.0 / (4 * k + 1) + z2 * (1.0 / (4 * k + 3));
  T x = z2 * z2;
#pragma unroll
  for (int n = NT / 2; n > 1; n = (n + 1) / 2) {
#pragma unroll
    for (int k = 0; k < n / 2; ++k) p[k] = p[2 * k] + x * p[2 * k + 1];
    if (n % 2) p[n / 2] = p[n - 1];
    if (n > 2) x = x * x;
  }
  return p[0];
}

// Terms the series needs for |z2| <= amax, to below rounding:
// 12 for kStepZ2 (0.04^12 / 25 < 1e-18), 6 for 2e-3 (5e-18).  ADAPT: one
// count for the whole wave (one vote and a scalar branch per series; the
// walk kernels, where a wave's lanes are 64 neighbouring pixels at one
// sample, so the choice does not depend on the launch shape); otherwise 12.
// A third, four-term rung (|z2| <= 1e-4) saves two multiply-adds per series
// but costs a second vote and more taken branches than it saves.
template <bool ADAPT> KG_INLINE int rc_terms(double amax) {
  return (!ADAPT || wave_any(!(amax <= 2.0e-3))) ? 12 : 6;
}

template <class T> KG_INLINE T rc_one_series(const T& z2, int nt = 12) {
  if (nt == 6) return rc_one_terms<6>(z2);
  return rc_one_terms<12>(z2);
}

// Jacobi state (sn, cn, dn), (E - F)/m and NP third-kind integrals at one
// argument.  T = double in the kernels; the op-counting CPU build
// instantiates it on its counting type.
template <class T, int NP> struct WalkState {
  T s, c, d;
  T D;
  T P[NP];
  T Lg[NP];  // ln|(1 + x)/(1 - x)| of a transformed integral (below)
};

// Per-pixel constants of a walk: the state at the step, the parameter and,
// per integral, n and n (m - n)(1 - n).
template <class T, int NP> struct WalkStep {
  WalkState<T, NP> b;
  T m;
  T n[NP], w2[NP];  // the walked characteristic and n (m - n)(1 - n)
  T pt[NP];         // sqrt((n - 1)(1 - m/n)) of a transformed integral, else 0
  T ipt2[NP];       // 1 / (2 pt)
  T cw[NP], ncw[NP];  // sqrt(-w2) and n / sqrt(-w2) where w2 < 0, else 0
  bool tr[NP];
  // some lane of the wave transforms an integral other than the first
  // (walk_init; the steps then walk every slot's log, else the first's only)
  bool tr_rest;
};

// Characteristics n > 1 (principal value) are walked through
//   Pi(n; am u | m) = -Pi(m/n; am u | m) + u + Lg(u) / (2 pt),
//   Lg(u) = ln|(1 + x)/(1 - x)|,  x = pt sn u / (cn u dn u),
//   pt = sqrt((n - 1)(1 - m/n))                           (DLMF 19.7.9)
// whose right-hand integral has 0 < m/n < 1 and no pole: Lg carries the
// singularity.  Lg walks too: for any real x0, x1
//   Lg(x1) - Lg(x0) = ln|(1 + y)/(1 - y)| = 2 y R_C(1, 1 - y^2),
//   y = (x1 - x0)/(1 - x0 x1),
// and with x = pt sn/(cn dn) the numerator is, by the subtraction theorem
// sn(a - b) (1 - m sn^2 a sn^2 b) = sn a cn b dn b - sn b cn a dn a,
//   y = pt sn v (1 - m sn^2(u+v) sn^2 u) / (cn(u+v) dn(u+v) cn u dn u - pt^2 sn(u+v) sn u),
// free of cancellation.  Where |y^2| > kStepZ2 (next to the pole) the lane
// takes Lg from its closed form instead.  Used for 0 <= m < 1, n > 1 and
// pt >= kTransformMinPt (1/pt amplifies rounding in Lg / (2 pt)); anywhere
// else n is walked as it is.
constexpr double kTransformMinPt = 0.05;

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
    w.ipt2[i] = w.tr[i] ? 0.5 / pt : T(0.0);
    w.n[i] = nw;
    w.w2[i] = nw * (m - nw) * (1.0 - nw);
    const bool circ = val(w.w2[i]) < 0.0;
    w.cw[i] = circ ? ksqrt(-w.w2[i]) : T(0.0);
    w.ncw[i] = circ ? nw / w.cw[i] : T(0.0);
  }
  bool rest = false;
#pragma unroll
  for (int i = 1; i < NP; ++i) rest = rest || w.tr[i];
  w.tr_rest = wave_any(rest);
}

// Direct entry into the walk at argument u: from Pi(n; am u), set the walked
// integral (and Lg) of slot i.  Next to the pole of the integrand the direct
// Pi(n) carries a rounding error of order eps / (1 - n sn^2), which the walk
// would carry to every later sample of the segment: there the walked value is
// set to NaN, so the next step is refused and evaluated directly too.
template <class T, int NP>
KG_INLINE void walk_enter(const WalkStep<T, NP>& w, int i, const T& Pn, const T& u, const T& sn, const T& cn,
                          const T& dn, WalkState<T, NP>& st) {
  if (w.tr[i]) {
    st.Lg[i] = log_ratio(w.pt[i] * sn, cn * dn);
    st.P[i] = -Pn + u + st.Lg[i] * w.ipt2[i];
    if (!(kfabs(val(st.Lg[i])) <= kLogGuard)) st.P[i] = T(NAN);
  } else {
    st.Lg[i] = T(0.0);
    st.P[i] = Pn;
    if (!(kfabs(val(1.0 - w.n[i] * (sn * sn))) >= kPoleGuard)) st.P[i] = T(NAN);
  }
}

// Pi(n; am u) of slot i from the walked state at argument u.
template <class T, int NP>
KG_INLINE T walk_pi(const WalkStep<T, NP>& w, int i, const WalkState<T, NP>& st, const T& u) {
  return w.tr[i] ? -st.P[i] + u + st.Lg[i] * w.ipt2[i] : st.P[i];
}

// 1/x_i for NP values with one reciprocal of their product.
template <class T, int NP> KG_INLINE void shared_reciprocals(const T (&x)[NP], T (&ix)[NP]) {
  if (NP == 1) {
    ix[0] = frcp(x[0]);
    return;
  }
  T prod = x[0];
#pragma unroll
  for (int i = 1; i < NP; ++i) prod = prod * x[i];
  const T ip = frcp(prod);
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    T o = ip;
#pragma unroll
    for (int k = 0; k < NP; ++k)
      if (k != i) o = o * x[k];
    ix[i] = o;
  }
}

// One step of a walked log ratio Lg = ln|(b + a)/(b - a)| from (a0, b0) to
// (a1, b1): Lg1 = Lg0 + ln|(1 + y)/(1 - y)|, y = (a1 b0 - a0 b1)/(b0 b1 - a0 a1),
// by the series while |y^2| <= kStepZ2, else from its closed form.  The
// caller passes 1/(b0 b1 - a0 a1) (walk_log_den), shared across integrals,
// and the series' term count.
template <class T> KG_INLINE T walk_log_y(const T& a0, const T& b0, const T& a1, const T& b1, const T& inv_den) {
  return (a1 * b0 - a0 * b1) * inv_den;
}

template <class T>
KG_INLINE T walk_log(const T& Lg0, const T& y, const T& a1, const T& b1, int nt) {
  const T y2 = y * y;
  T lg = Lg0 + 2.0 * y * rc_one_series(y2, nt);
  if (!(kfabs(val(y2)) <= kStepZ2)) lg = log_ratio(a1, b1);
  return lg;
}

template <class T> KG_INLINE T walk_log_den(const T& a0, const T& b0, const T& a1, const T& b1) {
  return b0 * b1 - a0 * a1;
}

// Steps of many samples at once (the anchor pass's step between segment
// starts) take the Pi increment in closed form beyond the series' range.
// With c = sqrt(|w2|) it is
//   w2 < 0:  (n / c) atan2(c q, p)      (arctan form, any step),
//   w2 > 0:  (n / c) artanh(c q / p)    (for z^2 up to kFarZ2).
// w2 < 0 only for n < 1, where 1 - n sn^2 > 0 throughout, and then q = 0
// forces p > 0 (q = 0 means sn u, sn v or sn(u+v) vanishes, and p reduces
// to 1 - n sn^2 of the other end): the atan2 branch cut (q = 0, p < 0) is
// never met, so the arctan form is continuous through p = 0 and needs no
// sign condition on p.
constexpr double kFarZ2 = 0.5;

// The Jacobi part of a step a -> a + du (DLMF 22.8.1-3): sn, cn, dn at
// a + du into out; false where one is not finite.
template <class T, int NP>
KG_INLINE bool walk_jacobi(const WalkState<T, NP>& a, const WalkStep<T, NP>& w, WalkState<T, NP>& out) {
  const WalkState<T, NP>& b = w.b;
  const T ss = a.s * b.s;
  const T inv = frcp(1.0 - w.m * ss * ss);
  const T S = (a.s * (b.c * b.d) + a.c * b.s * a.d) * inv;
  const T C = (a.c * b.c - ss * a.d * b.d) * inv;
  const T Dn = (a.d * b.d - w.m * ss * a.c * b.c) * inv;
  out.s = S;
  out.c = C;
  out.d = Dn;
  return kfinite(val(S)) & kfinite(val(C)) & kfinite(val(Dn));
}

// The rest of the step a -> a + du, given out.s, out.c, out.d from
// walk_jacobi (ok: its result): (E - F)/m, the NP third-kind integrals and
// their log terms.
template <class T, int NP, bool FAR = false, bool ADAPT = false>
KG_INLINE bool walk_integrals(const WalkState<T, NP>& a, const WalkStep<T, NP>& w, WalkState<T, NP>& out, bool ok) {
  const WalkState<T, NP>& b = w.b;
  const T ss = a.s * b.s;
  const T S = out.s;
  const T C = out.c;
  const T Dn = out.d;
  const T q = ss * S;
  out.D = a.D + b.D - q;
  const T cd = ss * C * Dn;
  // the NP values of p and one reciprocal of their product
  T pv[NP], r1v[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    r1v[i] = 1.0 - w.n[i] * (S * S);
    pv[i] = r1v[i] + w.n[i] * cd;
  }
  T ipv[NP];
  shared_reciprocals<T, NP>(pv, ipv);
  T z2v[NP];
  double zmax = 0.0;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    z2v[i] = w.w2[i] * (q * ipv[i]) * (q * ipv[i]);
    const double az = kfabs(val(z2v[i]));
    zmax = kmax(zmax, az <= kStepZ2 ? az : 0.0);  // others are refused
  }
  const int nt = rc_terms<ADAPT>(zmax);
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    // 1 - n sn^2 at both ends and p must share a sign: no pole of the
    // integrand (n > 1: principal value) lies inside the step
    const T r0 = 1.0 - w.n[i] * (a.s * a.s);
    const T& r1 = r1v[i];
    const T& p = pv[i];
    const T& ip = ipv[i];
    const T nq = w.n[i] * q * ip;
    const T& z2 = z2v[i];
    const bool series = kfabs(val(z2)) <= kStepZ2;
    const bool psign = (val(p) > 0.0) == (val(r1) > 0.0);
    // next to the integrand's pole (1 - n sn^2 -> 0: n > 1, or n -> 1 at
    // sn -> 1) 1 - n sn^2 carries an absolute rounding error of a few ulp,
    // which each step would add to Pi relative to 1 / (1 - n sn^2)
    const bool circ = FAR && val(w.w2[i]) < 0.0 && !(series && psign);
    // bitwise: every test is evaluated, with no branch per test
    ok = ok & (kfabs(val(r1)) >= kPoleGuard) & ((val(r0) > 0.0) == (val(r1) > 0.0)) &
         (circ | (psign & (series | (FAR && val(z2) <= kFarZ2))));
    T corr;
    if (circ) corr = w.ncw[i] * katan2(w.cw[i] * q, p);
    else corr = nq * ((!FAR || series) ? rc_one_series(z2, nt) : rc_1pe(-z2));
    out.P[i] = a.P[i] + b.P[i] + corr;
    ok = ok & kfinite(val(out.P[i]));
  }
  // the transformed integrals' log terms, one shared reciprocal; where no
  // lane of the wave transforms the second or third integral, the first's
  // alone (the same values: the general form multiplies by exact ones there)
  const T num = b.s * (1.0 - w.m * (S * S) * (a.s * a.s));
  const T cdn = C * Dn * (a.c * a.d);
  if (NP > 1 && !w.tr_rest) {
#pragma unroll
    for (int i = 1; i < NP; ++i) out.Lg[i] = T(0.0);
    if (w.tr[0]) {
      const T y = w.pt[0] * num * frcp(cdn - (w.pt[0] * w.pt[0]) * (S * a.s));
      const T y2 = y * y;
      const double ay = kfabs(val(y2));
      const int nty = rc_terms<ADAPT>(ay <= kStepZ2 ? ay : 0.0);
      T lg = a.Lg[0] + 2.0 * y * rc_one_series(y2, nty);
      if (!(kfabs(val(y2)) <= kStepZ2)) lg = log_ratio(w.pt[0] * S, C * Dn);
      out.Lg[0] = lg;
      ok = ok & (kfabs(val(lg)) <= kLogGuard);  // false for NaN
    } else {
      out.Lg[0] = T(0.0);
    }
    return ok & kfinite(val(out.D));
  }
  T yd[NP], iyd[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) yd[i] = w.tr[i] ? cdn - (w.pt[i] * w.pt[i]) * (S * a.s) : T(1.0);
  shared_reciprocals<T, NP>(yd, iyd);
  T yv[NP];
  double ymax = 0.0;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    yv[i] = w.pt[i] * num * iyd[i];
    const double ay = kfabs(val(yv[i] * yv[i]));
    ymax = kmax(ymax, (w.tr[i] & (ay <= kStepZ2)) ? ay : 0.0);
  }
  const int nty = rc_terms<ADAPT>(ymax);
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    if (w.tr[i]) {
      const T& y = yv[i];
      const T y2 = y * y;
      T lg = a.Lg[i] + 2.0 * y * rc_one_series(y2, nty);
      if (!(kfabs(val(y2)) <= kStepZ2)) lg = log_ratio(w.pt[i] * S, C * Dn);
      out.Lg[i] = lg;
      ok = ok & (kfabs(val(lg)) <= kLogGuard);  // false for NaN
    } else {
      out.Lg[i] = T(0.0);
    }
  }
  return ok & kfinite(val(out.D));
}

// One step a -> a + du.  Returns false (leaving `out` unspecified) when the
// step is not well conditioned; the caller then evaluates directly.  FAR
// admits z^2 < kFarZ2 (closed-form R_C) instead of |z^2| <= kStepZ2.
template <class T, int NP, bool INTEGRALS, bool FAR = false, bool ADAPT = false>
KG_INLINE bool walk_step(const WalkState<T, NP>& a, const WalkStep<T, NP>& w, WalkState<T, NP>& out) {
  const bool ok = walk_jacobi(a, w, out);
  if (!INTEGRALS) return ok;
  return walk_integrals<T, NP, FAR, ADAPT>(a, w, out, ok);
}

}  // namespace kg
