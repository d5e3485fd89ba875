// Per-pixel setup and per-sample evaluation of Kerr null geodesics in the
// Gralla-Lupsasca analytic formalism (GL19a arXiv:1910.12881 App. B,
// GL19b arXiv:1910.12873).
//
// Behaviour follows the reference (achael/kgeo):
//   conserved quantities         kerr_raytracing_ana.py:76-82
//   u_+/u_- and clamps           kerr_raytracing_utils.py:276-328
//   radial roots + rclass        kerr_raytracing_utils.py:330-380, 828-855
//   total Mino time              kerr_raytracing_utils.py:382-548
//   Mino-time grid               kerr_raytracing_ana.py:93-94 (numpy.linspace)
//   theta integrals              kerr_raytracing_ana.py:195-365
//   radial integrals + S/R       kerr_raytracing_ana.py:368-683
//   combination                  kerr_raytracing_ana.py:186-188
//
// Work split: setup() runs once per pixel and leaves everything that does
// not depend on tau in two small slot arrays (ThetaConsts, RadialConsts);
// the per-sample functions then cost one Jacobi evaluation and one shared
// Carlson base per modulus plus one R_J per third-kind integral.

#pragma once

#include "kg_math.hpp"

namespace kg {

constexpr double kMinSpin = 1.0e-4;  // kerr_raytracing_utils.py:9
constexpr double kEP = 1.0e-12;      // kerr_raytracing_utils.py:12

// Per-pixel status bits (uint32 status output).
enum : uint32_t {
  ST_NONFINITE = 1u << 0,        // some sample produced NaN/inf
  ST_DEGENERATE_ROOTS = 1u << 1, // resolvent root xi0 <= 0 or NaN (complex z)
  ST_MINSPIN_VORTICAL = 1u << 2, // |a|<MINSPIN with eta<=0 (reference raises)
  ST_LADDER_CAP = 1u << 3,       // AGM ladder hit its cap
  ST_LABEL_TIE = 1u << 4,        // rclass 0 (exact tie) or cube-root tie
};

// Ray-wide scalars.  For T = double they are computed on the host with the
// reference's numpy expressions and passed in; for tangent types they are
// rebuilt on the device from (a, th_o) by make_scalars().
template <class T> struct RayScalars {
  T a, a2, sin_th, cos_th, cos2_th, rplus, rminus;
  double r_o;
  double maxtaufrac;
  double h;          // +1 if th_o <= pi/2 else -1 (kerr_raytracing_ana.py:313)
  int minspin;       // |a| <  MINSPIN
  int spin_gt_min;   // |a| >  MINSPIN (kerr_raytracing_ana.py:529)
  int r_o_inf;
};

template <class T> KG_INLINE RayScalars<T> make_scalars(const T& a, const T& th_o, double r_o, double maxtaufrac) {
  RayScalars<T> c;
  c.a = a;
  c.a2 = a * a;
  c.sin_th = ksin(th_o);
  c.cos_th = kcos(th_o);
  c.cos2_th = c.cos_th * c.cos_th;
  T root = ksqrt(1.0 - c.a2);
  c.rplus = 1.0 + root;
  c.rminus = 1.0 - root;
  c.r_o = r_o;
  c.maxtaufrac = maxtaufrac;
  c.h = (val(th_o) <= kHalfPi) ? 1.0 : -1.0;
  c.minspin = kfabs(val(a)) < kMinSpin;
  c.spin_gt_min = kfabs(val(a)) > kMinSpin;
  c.r_o_inf = (r_o == INFINITY);
  return c;
}

// ---------------------------------------------------------------------------
// Pixel setup: lambda, eta, u_+-, roots, labels, tau_tot
// ---------------------------------------------------------------------------
template <class T> struct PixelSetup {
  T lam, eta, up, um, s_o;
  T rr[4], ri[4];  // real / imaginary parts of r1..r4
  T tau_tot;
  int rclass, thclass, rcase, status;
};

// u_+, u_- with the reference's clamps (kerr_raytracing_utils.py:276-328).
// For |a| < MINSPIN the reference aliases u_minus to u_plus, so every clamp
// applies to both; "alias" reproduces that.
template <class T>
KG_INLINE void uplus_uminus(const RayScalars<T>& c, const T& lam, const T& eta, T& up, T& um) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const bool alias = c.minspin;
  if (alias) {
    up = eta / (eta + lam * lam);
    um = up;
  } else {
    T dth = 0.5 * (1.0 - (eta + lam * lam) / c.a2);
    T disc = ksqrt(dth * dth + eta / c.a2);
    up = dth + disc;
    um = dth - disc;
  }
  if (val(c.cos2_th - up) > 0.0) { up = c.cos2_th; if (alias) um = up; }
  if (val(up * up) > 1.0) { up = T(1.0); if (alias) um = up; }
  if (val(um) == 0.0 && val(eta) < 0.0) { um = T(kEP); if (alias) up = um; }
  if (val(um) == 0.0 && val(eta) >= 0.0) { um = T(-kEP); if (alias) up = um; }
  if (val(eta) < 0.0 && val(c.cos2_th - um) < 0.0) { um = c.cos2_th; if (alias) up = um; }
}

// Radial roots in numpy's operation order (kerr_raytracing_utils.py:346-363).
// Notes on numpy semantics reproduced here:
//   * complex/real division is Smith's algorithm: (x+0j)/d == x * (1/d),
//     used for disc/108, th/3 and 0.25*B/z;
//   * the principal cube root of a conjugate pair sums to a real xi0;
//   * z = sqrt(xi0/2) is real whenever B != 0 (xi0 is then the largest
//     resolvent root, which is positive).
template <class T> KG_INLINE void radial_roots(const RayScalars<T>& c, const T& lam, const T& eta, PixelSetup<T>& px) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  T A = c.a2 - eta - lam * lam;
  T lma = lam - c.a;
  T B = 2.0 * (eta + lma * lma);
  T C = (-c.a2) * eta;
  T A2 = A * A;
  T P = -A2 / 12.0 - C;
  T Q = (-A / 3.0) * (A2 / 36.0 - C) - (B * B) / 8.0;
  T disc = 4.0 * (P * P * P) + 27.0 * (Q * Q);
  T d108 = disc * (1.0 / 108.0);
  T mQ = -0.5 * Q;
  T xi0;
  if (val(d108) >= 0.0) {
    T sd = ksqrt(d108);
    T wp = mQ + sd, wm = mQ - sd;
    T cp = kcbrt(kfabs(wp)), cm = kcbrt(kfabs(wm));
    T op = (val(wp) > 0.0) ? cp : (val(wp) < 0.0 ? -cp : T(0.0));
    T om = (val(wm) > 0.0) ? cm : (val(wm) < 0.0 ? -cm : T(0.0));
    xi0 = (op + om) - A / 3.0;
  } else if (val(d108) < 0.0) {
    // wp = mQ + i sd, wm = conj(wp); root with the largest real part
    T sd = ksqrt(-d108);
    T r = ksqrt(mQ * mQ + sd * sd);  // |w| (hypot)
    T th = katan2(sd, mQ);
    T th3 = th * (1.0 / 3.0);
    T cr = kcbrt(r);
    const T c23 = T(2.0943951023931953);  // 2j*pi/3 as Python evaluates it
    T x1 = cr * kcos(th3), x2 = cr * kcos(th3 + c23), x3 = cr * kcos(th3 - c23);
    T re;
    if (val(x1) > val(x2) && val(x1) > val(x3)) re = x1;
    else if (val(x2) > val(x1) && val(x2) > val(x3)) re = x2;
    else if (val(x3) > val(x1) && val(x3) > val(x2)) re = x3;
    else { re = T(NAN); px.status |= ST_LABEL_TIE; }
    xi0 = (re + re) - A / 3.0;
  } else {
    xi0 = T(NAN);
  }
  if (!(val(xi0) > 0.0)) px.status |= ST_DEGENERATE_ROOTS;
  T z = ksqrt(0.5 * xi0);
  T z2 = z * z;
  T t = (0.25 * B) * (1.0 / z);
  T base = -0.5 * A - z2;
  T wlo = base + t;
  T whi = base - t;
  // sqrt of a real-valued complex: (sqrt w, 0) for w >= 0, (0, sqrt(-w)) else
  const bool lo_c = !(val(wlo) >= 0.0);
  const bool hi_c = !(val(whi) >= 0.0);
  T slo = lo_c ? ksqrt(-wlo) : ksqrt(wlo);
  T shi = hi_c ? ksqrt(-whi) : ksqrt(whi);
  if (lo_c) {
    px.rr[0] = -z; px.ri[0] = -slo;
    px.rr[1] = -z; px.ri[1] = slo;
  } else {
    px.rr[0] = -z - slo; px.ri[0] = T(0.0);
    px.rr[1] = -z + slo; px.ri[1] = T(0.0);
  }
  if (hi_c) {
    px.rr[2] = z; px.ri[2] = -shi;
    px.rr[3] = z; px.ri[3] = shi;
  } else {
    px.rr[2] = z - shi; px.ri[2] = T(0.0);
    px.rr[3] = z + shi; px.ri[3] = T(0.0);
  }
  // NaN propagates as "imag != 0" in numpy's masks
  const bool im2 = lo_c;
  const bool im4 = hi_c;
  int rc = 0;
  if (im2) rc = 4;
  else if (im4) rc = 3;
  else if (val(px.rr[3]) < val(c.rplus)) rc = 2;
  else if (val(px.rr[3]) > val(c.rplus)) rc = 1;
  if (rc == 0) px.status |= ST_LABEL_TIE;
  px.rclass = rc;
  // r_integrate's case masks (kerr_raytracing_ana.py:405,474,555)
  px.rcase = im2 ? 4 : (im4 ? 3 : 2);
}

// Amplitude helpers for angles the reference builds with arcsin / arccos.
template <class T> KG_INLINE Amp<T> amp_from_sin(const T& x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  Amp<T> A;
  A.j = 0.0;
  A.s = x;
  A.c = ksqrt((1.0 - x) * (1.0 + x));
  return A;
}

template <class T> KG_INLINE Amp<T> amp_from_cos(const T& x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  // phi = arccos(x) in [0, pi]
  Amp<T> A;
  T s = ksqrt((1.0 - x) * (1.0 + x));
  if (val(x) >= 0.0) {
    A.j = 0.0; A.s = s; A.c = x;
  } else {
    A.j = 1.0; A.s = -s; A.c = -x;
  }
  return A;
}

// Reference special cases of ellip_pi_arr (scipy_ellip_binding.py:125-162)
// in the reference's order; phi is carried as the reduced amplitude, for
// which |phi| > pi/2 <=> j != 0.
template <class T>
KG_INLINE T pi_special(const LegendreBase<T>& L, const T& n, const T& m, bool& hit) {
  hit = true;
  const bool beyond = (L.j != 0.0);
  if (L.j == 0.0 && kfabs(val(L.s)) <= 1.0e-10) return L.s;   // |phi| <= 1e-10
  if (val(m) < -1.0e14) return T(0.0);
  if (val(L.y) < 0.0) return T(NAN);
  if (beyond && val(m) >= 1.0) return T(NAN);
  if (beyond && val(n) == 1.0) return T(INFINITY);
  if (val(1.0 - n * (L.s * L.s)) == 0.0) return T(INFINITY);
  if (kfabs(val(n)) == INFINITY) return T(0.0);
  hit = false;
  return T(0.0);
}

template <class T> KG_INLINE T ellip_pi_amp(const LegendreBase<T>& L, const T& n, const T& m, const T& Pic) {
  bool hit;
  T v = pi_special(L, n, m, hit);
  if (hit) return v;
  return legendre_Pi(L, n, Pic);
}

// F, E - F and NP third-kind integrals at one amplitude from one fused
// Carlson family (the per-sample path).  Special cases of ellip_pi_arr are
// applied per member exactly as ellip_pi_amp does.
template <class T, int NP> struct LegendreFam {
  T F, EmF;
  T Pi[NP];
};

template <class T, int NP>
KG_INLINE void legendre_family(const Amp<T>& A, const T& m, const T& Kc, const T& EmKc, const T (&n)[NP],
                               const T (&Pic)[NP], bool want_rd, LegendreFam<T, NP>& out) {
  LegendreBase<T> L;
  L.s = A.s;
  L.j = A.j;
  L.x = A.c * A.c;
  const T s2 = A.s * A.s;
  L.y = 1.0 - m * s2;
  T rho[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) rho[i] = 1.0 - n[i] * s2;
  CarlsonSet<T, NP> cs;
  carlson_family(L.x, L.y, rho, want_rd, cs);
  L.rf = cs.rf;
  out.F = periods(L.j, Kc) + L.s * cs.rf;
  out.EmF = want_rd ? periods(L.j, EmKc) - (m * (1.0 / 3.0)) * (s2 * L.s) * cs.rd : T(0.0);
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    bool hit;
    T sv = pi_special(L, n[i], m, hit);
    out.Pi[i] = hit ? sv : periods(L.j, Pic[i]) + L.s * cs.rf + (n[i] * (1.0 / 3.0)) * (s2 * L.s) * cs.rj[i];
  }
}

// complete Pi(n|m) with the reference's special cases at phi = pi/2
template <class T> KG_INLINE T ellip_pi_complete(const T& n, const T& m, const T& Kc) {
  if (val(m) < -1.0e14) return T(0.0);
  if (val(n) == 1.0) return T(INFINITY);
  if (kfabs(val(n)) == INFINITY) return T(0.0);
  return ellip_Pic(n, m, Kc);
}

// F(phi|m) for a reduced amplitude (scipy ellipkinc)
template <class T> KG_INLINE T ellip_F_amp(const LegendreBase<T>& L, const T& Kc) { return legendre_F(L, Kc); }

// ---------------------------------------------------------------------------
// Jacobi amplitude for a negative parameter k via m' = k/(k-1):
// sn(sqrt(1-m') x | k) = sqrt(1-m') sn(x|m')/dn(x|m'),
// cn(sqrt(1-m') x | k) = cn(x|m')/dn(x|m')  (kerr_raytracing_ana.py:262-277)
// ---------------------------------------------------------------------------
template <class T> struct NegAmp {
  T sinphi;  // sin(am(u|k)) unreduced, as the reference's sinPhi_tau
  T cosphi;  // cn(u|k)
  T dnk;     // dn(u|k) = 1/dn(x|m')
  Amp<T> A;  // reduced amplitude for the Legendre integrals
};

template <class T>
KG_INLINE NegAmp<T> neg_amplitude(const T& u, const T& sq, const JacobiLadder<T>& lad) {
  NegAmp<T> R;
  if (kfabs(val(u)) < kEP) {
    R.sinphi = u;
    R.cosphi = T(1.0);
    R.dnk = T(1.0);
    R.A.j = 0.0;
    R.A.s = u;
    R.A.c = T(1.0);
    return R;
  }
  T x = u / sq;
  T sn, cn, dn;
  jacobi_eval(lad, x, sn, cn, dn);
  T sphi = sq * sn / dn;
  T cphi = cn / dn;
  R.sinphi = sphi;
  R.cosphi = cphi;
  R.dnk = 1.0 / dn;
  R.A.j = krint(val(x) / (2.0 * val(lad.K)));
  const double sg = parity_sign(R.A.j);
  R.A.s = sg * sphi;
  R.A.c = sg * cphi;
  return R;
}

// ---------------------------------------------------------------------------
// theta part: per-pixel constants and per-sample evaluation
// ---------------------------------------------------------------------------
enum ThSlot {
  TH_w, TH_sG, TH_pref, TH_sGph, TH_sGt, TH_k, TH_mp, TH_sq, TH_Kk, TH_EmKk, TH_Pic, TH_n,
  TH_A, TH_B, TH_C, TH_h, TH_NSLOT
};
// kinds
enum : int { THK_NONE = 0, THK_ORD = 1, THK_ORD_K0 = 2, THK_VORT = 3 };

template <class T> struct ThetaConsts {
  T v[TH_NSLOT];
  int kind;
  int eta_zero;  // eta == 0: ordinary branch only if some eta>0 in the image
};

template <class T>
KG_INLINE void theta_consts(const RayScalars<T>& c, const PixelSetup<T>& px, bool do_phi_t, ThetaConsts<T>& tc) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  for (int i = 0; i < TH_NSLOT; ++i) tc.v[i] = T(0.0);
  tc.kind = THK_NONE;
  tc.eta_zero = (val(px.eta) == 0.0);
  const T& up = px.up;
  const T& um = px.um;
  const T& s = px.s_o;
  if (val(px.eta) >= 0.0) {
    // ordinary motion (kerr_raytracing_ana.py:217-296)
    T k, a2um;
    if (c.minspin) {
      k = T(0.0);
      a2um = -(px.eta + px.lam * px.lam);
    } else {
      k = up / um;
      a2um = c.a2 * um;
    }
    T w = ksqrt(-a2um);
    T pref = 1.0 / w;
    Amp<T> Ao = amp_from_sin(c.cos_th / ksqrt(up));
    LegendreBase<T> Lo = legendre_base(Ao, k);
    T Kk = ellip_K(k);
    T Fo = ellip_F_amp(Lo, Kk);
    T Gth_o = -pref * Fo;
    tc.v[TH_w] = w;
    tc.v[TH_sG] = s * Gth_o;
    tc.v[TH_pref] = pref;
    tc.v[TH_k] = k;
    T mp = k / (k - 1.0);
    tc.v[TH_mp] = mp;
    tc.v[TH_sq] = ksqrt(1.0 - mp);
    tc.v[TH_n] = up;
    tc.v[TH_A] = (-s) * ksqrt(up);  // theta = arccos(A sinPhi)
    tc.v[TH_B] = 2.0 * up * pref;   // G_t prefactor
    tc.v[TH_C] = -tc.v[TH_B] / (2.0 * k);  // G_t = TH_C (E - F) - sGt for k != 0
    tc.kind = (val(k) == 0.0) ? THK_ORD_K0 : THK_ORD;
    if (do_phi_t) {
      T EmKk = ellip_EmK(k);
      T Pic = ellip_pi_complete(up, k, Kk);
      tc.v[TH_Kk] = Kk;
      tc.v[TH_EmKk] = EmKk;
      tc.v[TH_Pic] = Pic;
      T Gph_o = -pref * ellip_pi_amp(Lo, up, k, Pic);
      tc.v[TH_sGph] = s * Gph_o;
      T Gt_o;
      if (val(k) == 0.0) {
        T arg = kasin(c.cos_th / ksqrt(up));
        Gt_o = 2.0 * up * pref * 0.125 * (-2.0 * arg + ksin(2.0 * arg));
      } else {
        T EmF = legendre_EmF(Lo, k, EmKk);
        Gt_o = 2.0 * up * pref * EmF / (2.0 * k);
      }
      tc.v[TH_sGt] = s * Gt_o;
    }
  } else if (val(px.eta) < 0.0) {
    // vortical motion (kerr_raytracing_ana.py:300-363)
    T a2um = c.a2 * um;
    T prefA = 1.0 / ksqrt(a2um);
    T prefB = prefA / (1.0 - um);
    T prefC = ksqrt(um / c.a2);
    Amp<T> Ao = amp_from_sin(ksqrt((c.cos2_th - um) / (up - um)));
    T k = 1.0 - up / um;
    T upm = (up - um) / (1.0 - um);
    LegendreBase<T> Lo = legendre_base(Ao, k);
    T Kk = ellip_K(k);
    T Gth_o = -c.h * prefA * ellip_F_amp(Lo, Kk);
    tc.v[TH_w] = ksqrt(a2um);
    tc.v[TH_sG] = s * Gth_o;
    tc.v[TH_pref] = prefB;
    tc.v[TH_k] = k;
    T mp = k / (k - 1.0);
    tc.v[TH_mp] = mp;
    tc.v[TH_sq] = ksqrt(1.0 - mp);
    tc.v[TH_n] = upm;
    tc.v[TH_A] = um;
    tc.v[TH_B] = up - um;
    tc.v[TH_C] = prefC;
    tc.v[TH_h] = T(c.h);
    tc.kind = THK_VORT;
    if (do_phi_t) {
      T EmKk = ellip_EmK(k);
      T Pic = ellip_pi_complete(upm, k, Kk);
      tc.v[TH_Kk] = Kk;
      tc.v[TH_EmKk] = EmKk;
      tc.v[TH_Pic] = Pic;
      T Gph_o = -c.h * prefB * ellip_pi_amp(Lo, upm, k, Pic);
      T Eo = legendre_F(Lo, Kk) + legendre_EmF(Lo, k, EmKk);
      T Gt_o = -c.h * prefC * Eo;
      tc.v[TH_sGph] = s * Gph_o;
      tc.v[TH_sGt] = s * Gt_o;
    }
  }
}

// State of the theta integrals at one Jacobi argument u (= snarg): the
// Jacobi functions of parameter k at u, and the Legendre integrals at
// amplitude am(u|k).  The sample loops either evaluate it directly
// (theta_state) or advance it along the Mino-time grid with the addition
// theorems (kg_step.hpp); theta_out turns it into theta, G_phi, G_t.
template <class T> struct ThState {
  T s, c, d;   // sn, cn, dn (u | k): s is the reference's sinPhi_tau
  T F, EmF;    // F(am u | k) (= u) and E - F
  T Pi;        // Pi(n; am u | k)
};

// KIND = THK_* fixes the motion class at compile time (the kernels' loops);
// KIND = -1 reads it from tc.kind.
template <class T, int KIND = -1>
KG_INLINE void theta_state(const ThetaConsts<T>& tc, const JacobiLadder<T>& lad, const T& snarg, bool do_phi_t,
                           ThState<T>& st) {
  const T* v = tc.v;
  const int kind = (KIND >= 0) ? KIND : tc.kind;
  NegAmp<T> R = neg_amplitude(snarg, v[TH_sq], lad);
  st.s = R.sinphi;
  st.c = R.cosphi;
  st.d = R.dnk;
  if (!do_phi_t) return;
  const T& k = v[TH_k];
  LegendreFam<T, 1> fam;
  const T nn[1] = {v[TH_n]};
  const T pc[1] = {v[TH_Pic]};
  legendre_family(R.A, k, v[TH_Kk], v[TH_EmKk], nn, pc, kind != THK_ORD_K0, fam);
  st.F = fam.F;
  st.EmF = fam.EmF;
  st.Pi = fam.Pi[0];
}

template <class T, int KIND = -1>
KG_INLINE void theta_out(const ThetaConsts<T>& tc, const T& snarg, const ThState<T>& st, bool do_phi_t, T& th, T& Gph,
                         T& Gt) {
  const T* v = tc.v;
  const int kind = (KIND >= 0) ? KIND : tc.kind;
  if (kind == THK_VORT) {
    th = acos_poly(v[TH_h] * ksqrt(v[TH_A] + v[TH_B] * (st.s * st.s)));
  } else {
    th = acos_poly(v[TH_A] * st.s);
  }
  if (!do_phi_t) return;
  Gph = v[TH_pref] * st.Pi - v[TH_sGph];
  if (kind == THK_ORD) {
    Gt = v[TH_C] * st.EmF - v[TH_sGt];
  } else if (kind == THK_ORD_K0) {
    // k == 0: m' = 0, am(u|0) = u (kerr_raytracing_ana.py:290-291), so
    // sn = sin Phi, cn = cos Phi and sin 2 Phi = 2 sn cn
    T Phi = snarg;
    Gt = -v[TH_B] * 0.125 * (-2.0 * Phi + 2.0 * (st.s * st.c)) - v[TH_sGt];
  } else {
    Gt = v[TH_C] * (st.F + st.EmF) - v[TH_sGt];
  }
}

template <class T> KG_INLINE T theta_arg(const ThetaConsts<T>& tc, const T& tau) {
  return tc.v[TH_w] * (-tau + tc.v[TH_sG]);
}

template <class T, int KIND = -1>
KG_INLINE void theta_eval(const ThetaConsts<T>& tc, const JacobiLadder<T>& lad, const T& tau, bool do_phi_t, T& th,
                          T& Gph, T& Gt) {
  T snarg = theta_arg(tc, tau);
  ThState<T> st;
  theta_state<T, KIND>(tc, lad, snarg, do_phi_t, st);
  theta_out<T, KIND>(tc, snarg, st, do_phi_t, th, Gph, Gt);
}

// ---------------------------------------------------------------------------
// radial part
// ---------------------------------------------------------------------------
// Shared slots
enum RSlot {
  // shared: modulus, K(k), E(k)-K(k); observer argument u0 = F(aux|k) and
  // sn, cn, dn there; X = u0 - Xs*tau
  R_k, R_K, R_EmK, R_U0, R_Xs, R_sn0, R_cn0, R_dn0,
  // case 2 (kerr_raytracing_ana.py:554-624)
  C2_rr1 = 8, C2_rr2, C2_rr3, C2_rr4, C2_rr31, C2_rr41, C2_rr43, C2_rrp3, C2_rrm3,
  C2_n1, C2_np, C2_nm, C2_Pc1, C2_Pcp, C2_Pcm, C2_Eb, C2_Pb1, C2_Pbp, C2_Pbm,
  C2_c, C2_fp, C2_fm, C2_sq, C2_dX, C2_drso, C2_ro3, C2_h12,
  C2_irp3, C2_irm3, C2_Ho,  // 1/(r+ - r3), 1/(r- - r3), H at the observer
  // case 3 (kerr_raytracing_ana.py:473-552)
  C3_AA = 8, C3_BB, C3_rr1, C3_rr2, C3_al0, C3_alp, C3_alm, C3_nn0, C3_nnp, C3_nnm,
  C3_Pc0, C3_Pcp, C3_Pcm, C3_R1b0, C3_R2b0, C3_R1bp, C3_R1bm, C3_c0, C3_cp, C3_cm,
  C3_pr, C3_sqAB, C3_Ipd, C3_Imd, C3_hasm,
  C3_p0, C3_pp, C3_pm,  // p1 of the three R1 blocks
  C3_i0, C3_ip, C3_im,  // 1/(1 - al^2) of the three R1 blocks
  C3_iIpd, C3_iImd,     // 1/Ipd, 1/Imd
  C3_r2a, C3_r2b, C3_r2c,  // R2 block: al^2/den, 1/(al^2 - 1), (2k - n)/den
  // case 4 (kerr_raytracing_ana.py:404-471)
  C4_a2 = 8, C4_b1, C4_g0, C4_gp, C4_gm, C4_n0, C4_np, C4_nm, C4_Pc0, C4_Pcp, C4_Pcm,
  C4_S1b0, C4_S2b0, C4_S1bp, C4_S1bm, C4_P1f, C4_P2f, C4_Ppf, C4_Pmf, C4_aa1, C4_Ipf, C4_Imf,
  C4_q0, C4_qp, C4_qm,  // p2 of the three S1 blocks
  C4_l0, C4_lp, C4_lm,  // ln|(1 - p2)/(1 + p2)| of the three S1 blocks
  C4_i0, C4_ip, C4_im,  // 1/(1 + al^2) of the three S1 blocks
  R_NSLOT = 44
};

template <class T> struct RadialConsts {
  T v[R_NSLOT];
  int rcase;
};

// R1/R2 building blocks (kerr_raytracing_ana.py:666-683), given sin, cos of
// the (unreduced) amplitude, sqrt(1 - j sin^2), and the shared Legendre base.
template <class T>
KG_INLINE T R1_block(const T& al, const T& Pi, const T& sphi, const T& s2phi, const T& j) {
  T al2 = al * al;
  T p1 = ksqrt((al2 - 1.0) / (j + (1.0 - j) * al2));
  T f1 = 0.5 * p1 * klog(kfabs((p1 * s2phi + sphi) / (p1 * s2phi - sphi)));
  return (Pi - al * f1) / (1.0 - al2);
}

template <class T>
KG_INLINE T R2_block(const T& al, const T& nn, const T& R1, const T& F, const T& E, const T& sphi,
                     const T& cphi, const T& s2phi, const T& j) {
  T al2 = al * al;
  T den = j + (1.0 - j) * al2;
  T R2 = (F - (al2 / den) * (E - al * sphi * s2phi / (1.0 + al * cphi))) / (al2 - 1.0);
  return R2 + (2.0 * j - nn) * R1 / den;
}

// S1/S2 building blocks (kerr_raytracing_ana.py:638-663).  tan(phi) enters
// only through (al - tan)/(1 + al tan) = (al c - s)/(c + al s).
// The R1 and S1 blocks with their per-pixel p and the log of their f term
// supplied: lg = ln|(p1 dn + sn)/(p1 dn - sn)| for R1 and
// lg = ln|(1 + p2 dn)/(1 - p2 dn)|, lc = ln|(1 - p2)/(1 + p2)| for S1 (the
// f1, f2 of R1_block and S1_block).  The sample loops take lg from a direct
// log or from its walk.
template <class T> KG_INLINE T R1_from(const T& al, const T& Pi, const T& p1, const T& lg, const T& inv1mal2) {
  return (Pi - al * (0.5 * p1 * lg)) * inv1mal2;
}
template <class T>
KG_INLINE T S1_from(const T& al, const T& Pi, const T& F, const T& p2, const T& lg, const T& lc, const T& inv1pal2) {
  return (F + (al * al) * Pi - al * (0.5 * p2 * (lg + lc))) * inv1pal2;
}
template <class T> KG_INLINE T R1_p(const T& al, const T& j) {
  T al2 = al * al;
  return ksqrt((al2 - 1.0) / (j + (1.0 - j) * al2));
}
template <class T> KG_INLINE T S1_p(const T& al, const T& j) {
  T al2 = al * al;
  return ksqrt((1.0 + al2) / (1.0 - j + al2));
}

template <class T>
KG_INLINE T S1_block(const T& al, const T& Pi, const T& F, const T& s2phi, const T& j) {
  T al2 = al * al;
  T al2p1 = 1.0 + al2;
  T al2j = 1.0 - j + al2;
  T p2 = ksqrt(al2p1 / al2j);
  T f2 = 0.5 * p2 * klog(kfabs(((1.0 - p2) * (1.0 + p2 * s2phi)) / ((1.0 + p2) * (1.0 - p2 * s2phi))));
  return (F + al2 * Pi - al * f2) / al2p1;
}

template <class T>
KG_INLINE T S2_block(const T& al, const T& S1, const T& F, const T& E, const T& sphi, const T& cphi,
                     const T& s2phi, const T& j) {
  T al2 = al * al;
  T al2p1 = 1.0 + al2;
  T al2j = 1.0 - j + al2;
  T S2a = (1.0 - j) * F + al2 * E + al2 * s2phi * (al * cphi - sphi) / (cphi + al * sphi) - al2 * al;
  S2a = -S2a / (al2p1 * al2j);
  T S2b = S1 * (1.0 / al2p1 + (1.0 - j) / al2j);
  return S2a + S2b;
}

// sin, cos of an angle with |angle| < 3pi/2 -> reduced amplitude
template <class T> KG_INLINE Amp<T> amp_from_sc(const T& sb, const T& cb) {
  Amp<T> A;
  if (val(cb) >= 0.0) {
    A.j = 0.0; A.s = sb; A.c = cb;
  } else {
    A.j = (val(sb) >= 0.0) ? 1.0 : -1.0;
    A.s = -sb; A.c = -cb;
  }
  return A;
}

// sin and cos of atan(x) + atan(g) without forming the angle
// (the case-4 amplitudes, kerr_raytracing_ana.py:432, utils.py:437-438);
// x = +inf gives the r_o = infinity limit.
template <class T> KG_INLINE void atan_sum_sc(const T& x, const T& g, bool x_inf, T& sb, T& cb) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  T dg = ksqrt(1.0 + g * g);
  if (x_inf) {
    sb = 1.0 / dg;
    cb = -g / dg;
    return;
  }
  T d = ksqrt(1.0 + x * x) * dg;
  sb = (x + g) / d;
  cb = (1.0 - x * g) / d;
}

template <class T>
KG_INLINE void radial_consts(const RayScalars<T>& c, const PixelSetup<T>& px, bool do_phi_t, RadialConsts<T>& rc) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  for (int i = 0; i < R_NSLOT; ++i) rc.v[i] = T(0.0);
  rc.rcase = px.rcase;
  T* v = rc.v;
  const T& rplus = c.rplus;
  const T& rminus = c.rminus;
  const double r_o = c.r_o;
  if (px.rcase == 4) {
    T a1 = px.ri[3], b1 = px.rr[3], a2 = px.ri[1], b2 = px.rr[1];
    T CC = ksqrt(sq(a1 - a2) + sq(b1 - b2));
    T DD = ksqrt(sq(a1 + a2) + sq(b1 - b2));
    T k4 = (4.0 * CC * DD) / sq(CC + DD);
    T x4rp = (rplus + b1) / a2;
    T x4rm = (rminus + b1) / a2;
    T g0 = ksqrt((4.0 * a2 * a2 - sq(CC - DD)) / (sq(CC + DD) - 4.0 * a2 * a2));
    T gp = (g0 * x4rp - 1.0) / (g0 + x4rp);
    T gm = (g0 * x4rm - 1.0) / (g0 + x4rm);
    T sb, cb;
    atan_sum_sc(c.r_o_inf ? T(0.0) : (r_o + b1) / a2, g0, c.r_o_inf, sb, cb);
    T pref = 2.0 / (CC + DD);
    T K = ellip_K(k4);
    Amp<T> Ab = amp_from_sc(sb, cb);
    LegendreBase<T> Lb = legendre_base(Ab, k4);
    T Fb = legendre_F(Lb, K);
    T s2b = ksqrt(1.0 - k4 * (sb * sb));
    v[R_k] = k4;
    v[R_K] = K;
    v[R_Xs] = 1.0 / pref;
    v[R_U0] = Fb;
    v[R_sn0] = sb;
    v[R_cn0] = cb;
    v[R_dn0] = s2b;
    v[C4_a2] = a2;
    v[C4_b1] = b1;
    v[C4_g0] = g0;
    if (do_phi_t) {
      T EmK = ellip_EmK(k4);
      v[R_EmK] = EmK;
      T n0 = 1.0 + g0 * g0, np = 1.0 + gp * gp, nm = 1.0 + gm * gm;
      T Pc0 = ellip_pi_complete(n0, k4, K), Pcp = ellip_pi_complete(np, k4, K), Pcm = ellip_pi_complete(nm, k4, K);
      v[C4_gp] = gp; v[C4_gm] = gm;
      v[C4_n0] = n0; v[C4_np] = np; v[C4_nm] = nm;
      v[C4_Pc0] = Pc0; v[C4_Pcp] = Pcp; v[C4_Pcm] = Pcm;
      T Eb = Fb + legendre_EmF(Lb, k4, EmK);
      T S1b0 = S1_block(g0, ellip_pi_amp(Lb, n0, k4, Pc0), Fb, s2b, k4);
      v[C4_S1b0] = S1b0;
      v[C4_S2b0] = S2_block(g0, S1b0, Fb, Eb, sb, cb, s2b, k4);
      v[C4_S1bp] = S1_block(gp, ellip_pi_amp(Lb, np, k4, Pcp), Fb, s2b, k4);
      v[C4_S1bm] = S1_block(gm, ellip_pi_amp(Lb, nm, k4, Pcm), Fb, s2b, k4);
      const T q0 = S1_p(g0, k4), qp = S1_p(gp, k4), qm = S1_p(gm, k4);
      v[C4_q0] = q0; v[C4_qp] = qp; v[C4_qm] = qm;
      v[C4_l0] = klog(kfabs((1.0 - q0) / (1.0 + q0)));
      v[C4_lp] = klog(kfabs((1.0 - qp) / (1.0 + qp)));
      v[C4_lm] = klog(kfabs((1.0 - qm) / (1.0 + qm)));
      v[C4_i0] = 1.0 / (1.0 + g0 * g0);
      v[C4_ip] = 1.0 / (1.0 + gp * gp);
      v[C4_im] = 1.0 / (1.0 + gm * gm);
      T aa0 = (a2 / g0) * (1.0 + g0 * g0);
      v[C4_P1f] = pref * aa0;
      v[C4_P2f] = pref * (aa0 * aa0);
      v[C4_Ppf] = pref * (1.0 + g0 * g0) / (g0 * g0 + g0 * x4rp);
      v[C4_Pmf] = pref * (1.0 + g0 * g0) / (g0 * g0 + g0 * x4rm);
      v[C4_aa1] = (a2 / g0 - b1);
      v[C4_Ipf] = g0 / (a2 - a2 * g0 * x4rp);
      v[C4_Imf] = g0 / (a2 - a2 * g0 * x4rm);
    }
  } else if (px.rcase == 3) {
    T a1 = px.ri[3], b1 = px.rr[3];
    T rr1 = px.rr[0], rr2 = px.rr[1];
    T rr21 = rr2 - rr1;
    T rrp1 = rplus - rr1, rrp2 = rplus - rr2;
    T rrm1 = rminus - rr1, rrm2 = rminus - rr2;
    T AA = ksqrt(a1 * a1 + sq(b1 - rr2));
    T BB = ksqrt(a1 * a1 + sq(b1 - rr1));
    T k3 = (sq(AA + BB) - rr21 * rr21) / (4.0 * AA * BB);
    T x3ro;
    if (c.r_o_inf) {
      x3ro = (AA - BB) / (AA + BB);
    } else {
      T rro1 = r_o - rr1, rro2 = r_o - rr2;
      x3ro = (AA * rro1 - BB * rro2) / (AA * rro1 + BB * rro2);
    }
    T pref = 1.0 / ksqrt(AA * BB);
    T K = ellip_K(k3);
    T sb = ksqrt((1.0 - x3ro) * (1.0 + x3ro));  // sin(arccos x3ro)
    T cb = x3ro;
    Amp<T> Ab = amp_from_sc(sb, cb);
    LegendreBase<T> Lb = legendre_base(Ab, k3);
    T Fb = legendre_F(Lb, K);
    T s2b = ksqrt(1.0 - k3 * (sb * sb));
    v[R_k] = k3;
    v[R_K] = K;
    v[R_Xs] = 1.0 / pref;
    v[R_U0] = Fb;
    v[R_sn0] = sb;
    v[R_cn0] = cb;
    v[R_dn0] = s2b;
    v[C3_AA] = AA;
    v[C3_BB] = BB;
    v[C3_rr1] = rr1;
    v[C3_rr2] = rr2;
    if (do_phi_t) {
      T x3rp = (AA * rrp1 - BB * rrp2) / (AA * rrp1 + BB * rrp2);
      T x3rm = (AA * rrm1 - BB * rrm2) / (AA * rrm1 + BB * rrm2);
      T alp = -1.0 / x3rp;
      T alm = -1.0 / x3rm;
      T al0 = (BB + AA) / (BB - AA);
      T EmK = ellip_EmK(k3);
      v[R_EmK] = EmK;
      T nn0 = (al0 * al0) / (al0 * al0 - 1.0);
      T nnp = (alp * alp) / (alp * alp - 1.0);
      T nnm = (alm * alm) / (alm * alm - 1.0);
      T Pc0 = ellip_pi_complete(nn0, k3, K), Pcp = ellip_pi_complete(nnp, k3, K);
      T Pcm = c.spin_gt_min ? ellip_pi_complete(nnm, k3, K) : T(0.0);
      v[C3_al0] = al0; v[C3_alp] = alp; v[C3_alm] = alm;
      v[C3_nn0] = nn0; v[C3_nnp] = nnp; v[C3_nnm] = nnm;
      v[C3_Pc0] = Pc0; v[C3_Pcp] = Pcp; v[C3_Pcm] = Pcm;
      T Eb = Fb + legendre_EmF(Lb, k3, EmK);
      T R1b0 = R1_block(al0, ellip_pi_amp(Lb, nn0, k3, Pc0), sb, s2b, k3);
      v[C3_R1b0] = R1b0;
      v[C3_R2b0] = R2_block(al0, nn0, R1b0, Fb, Eb, sb, cb, s2b, k3);
      v[C3_R1bp] = R1_block(alp, ellip_pi_amp(Lb, nnp, k3, Pcp), sb, s2b, k3);
      v[C3_R1bm] = c.spin_gt_min ? R1_block(alm, ellip_pi_amp(Lb, nnm, k3, Pcm), sb, s2b, k3) : T(0.0);
      v[C3_hasm] = c.spin_gt_min ? T(1.0) : T(0.0);
      v[C3_p0] = R1_p(al0, k3);
      v[C3_pp] = R1_p(alp, k3);
      v[C3_pm] = R1_p(alm, k3);
      v[C3_i0] = 1.0 / (1.0 - al0 * al0);
      v[C3_ip] = 1.0 / (1.0 - alp * alp);
      v[C3_im] = 1.0 / (1.0 - alm * alm);
      T sqAB = ksqrt(AA * BB);
      T c0 = (2.0 * rr21 * sqAB) / (BB * BB - AA * AA);
      v[C3_c0] = c0;
      v[C3_cp] = (2.0 * rr21 * sqAB) / (BB * rrp2 - AA * rrp1);
      v[C3_cm] = (2.0 * rr21 * sqAB) / (BB * rrm2 - AA * rrm1);
      v[C3_pr] = (BB * rr2 + AA * rr1) / (BB + AA);
      v[C3_sqAB] = sqAB;
      v[C3_Ipd] = BB * rrp2 + AA * rrp1;
      v[C3_Imd] = BB * rrm2 + AA * rrm1;
      v[C3_iIpd] = 1.0 / v[C3_Ipd];
      v[C3_iImd] = 1.0 / v[C3_Imd];
      {
        const T al2 = al0 * al0;
        const T den = k3 + (1.0 - k3) * al2;
        v[C3_r2a] = al2 / den;
        v[C3_r2b] = 1.0 / (al2 - 1.0);
        v[C3_r2c] = (2.0 * k3 - nn0) / den;
      }
    }
  } else {
    T rr1 = px.rr[0], rr2 = px.rr[1], rr3 = px.rr[2], rr4 = px.rr[3];
    T rr31 = rr3 - rr1, rr32 = rr3 - rr2, rr41 = rr4 - rr1, rr42 = rr4 - rr2, rr43 = rr4 - rr3;
    T rrp3 = rplus - rr3, rrm3 = rminus - rr3, rrp4 = rplus - rr4, rrm4 = rminus - rr4;
    T k2 = (rr32 * rr41) / (rr31 * rr42);
    T x2ro = c.r_o_inf ? ksqrt(rr31 / rr41) : ksqrt((rr31 * (r_o - rr4)) / (rr41 * (r_o - rr3)));
    T K = ellip_K(k2);
    T sb = x2ro;
    T cb = ksqrt((1.0 - x2ro) * (1.0 + x2ro));
    Amp<T> Ab = amp_from_sc(sb, cb);
    LegendreBase<T> Lb = legendre_base(Ab, k2);
    T Fb = legendre_F(Lb, K);
    T sq3142 = ksqrt(rr31 * rr42);
    v[R_k] = k2;
    v[R_K] = K;
    v[R_Xs] = 0.5 * sq3142;
    v[R_U0] = Fb;
    v[R_sn0] = sb;
    v[R_cn0] = cb;
    v[R_dn0] = ksqrt(1.0 - k2 * (sb * sb));
    v[C2_rr1] = rr1; v[C2_rr2] = rr2; v[C2_rr3] = rr3; v[C2_rr4] = rr4;
    v[C2_rr31] = rr31; v[C2_rr41] = rr41; v[C2_rr43] = rr43;
    if (do_phi_t) {
      T EmK = ellip_EmK(k2);
      v[R_EmK] = EmK;
      v[C2_rrp3] = rrp3; v[C2_rrm3] = rrm3;
      T n1 = rr41 / rr31;
      T np = (rrp3 * rr41) / (rrp4 * rr31);
      T nm = (rrm3 * rr41) / (rrm4 * rr31);
      v[C2_n1] = n1; v[C2_np] = np; v[C2_nm] = nm;
      T Pc1 = ellip_pi_complete(n1, k2, K), Pcp = ellip_pi_complete(np, k2, K), Pcm = ellip_pi_complete(nm, k2, K);
      v[C2_Pc1] = Pc1; v[C2_Pcp] = Pcp; v[C2_Pcm] = Pcm;
      v[C2_Eb] = Fb + legendre_EmF(Lb, k2, EmK);
      v[C2_Pb1] = ellip_pi_amp(Lb, n1, k2, Pc1);
      v[C2_Pbp] = ellip_pi_amp(Lb, np, k2, Pcp);
      v[C2_Pbm] = ellip_pi_amp(Lb, nm, k2, Pcm);
      T cfac = 2.0 / sq3142;
      v[C2_c] = cfac;
      v[C2_fp] = cfac * (rr43 / (rrp3 * rrp4));
      v[C2_fm] = cfac * (rr43 / (rrm3 * rrm4));
      v[C2_sq] = sq3142;
      v[C2_dX] = -0.5 * sq3142;
      T Rpot_o = (r_o - rr1) * (r_o - rr2) * (r_o - rr3) * (r_o - rr4);
      v[C2_drso] = ksqrt(Rpot_o);
      v[C2_ro3] = r_o - rr3;
      v[C2_h12] = 0.5 * (rr1 * rr4 + rr2 * rr3);
      v[C2_irp3] = 1.0 / rrp3;
      v[C2_irm3] = 1.0 / rrm3;
      v[C2_Ho] = v[C2_drso] / v[C2_ro3];
    }
  }
}

template <class T> struct RadialOut {
  T r, Ip, Im, I1, I2;
};

// State of the radial integrals at one Jacobi argument X: sn, cn, dn (X | k)
// and the Legendre integrals at amplitude am(X | k) (F = X, E - F and the
// case's three third-kind integrals).  radial_state evaluates it directly;
// kg_step.hpp advances it along the grid; radial_out turns it into r and
// I_+, I_-, I_1, I_2.
template <class T> struct RState {
  T s, c, d;
  T F, EmF;
  T Pi[3];
  T lg[3];  // cases 3 and 4: the logs of the three R1 / S1 blocks
};

// The block logs of cases 3 and 4 at one Jacobi state, as log_ratio(a, b) =
// ln|(b + a)/(b - a)|: R1 a = sn, b = p1 dn; S1 a = p2 dn, b = 1.
template <class T, int CASE>
KG_INLINE void block_log_args(const RadialConsts<T>& rc, int i, const T& sn, const T& dn, T& a, T& b) {
  const T* v = rc.v;
  if (CASE == 3) {
    a = sn;
    b = v[C3_p0 + i] * dn;
  } else {
    a = v[C4_q0 + i] * dn;
    b = T(1.0);
  }
}

template <class T> KG_INLINE T log_ratio(const T& a, const T& b) {
  const T arg = (kfabs(val(a)) < kfabs(val(b))) ? 2.0 * a / (b - a) : 2.0 * b / (a - b);
  return klog1p(arg);
}

template <class T> KG_INLINE T radial_arg(const RadialConsts<T>& rc, const T& tau) {
  return rc.v[R_U0] - rc.v[R_Xs] * tau;
}

// The case's Pi parameters n_i and complete integrals, in Legendre-family order.
template <class T, int CASE>
KG_INLINE void radial_pi_params(const RadialConsts<T>& rc, T (&nn)[3], T (&pc)[3]) {
  const T* v = rc.v;
  const int rcase = (CASE >= 0) ? CASE : rc.rcase;
  if (rcase == 4) {
    nn[0] = v[C4_n0]; nn[1] = v[C4_np]; nn[2] = v[C4_nm];
    pc[0] = v[C4_Pc0]; pc[1] = v[C4_Pcp]; pc[2] = v[C4_Pcm];
  } else if (rcase == 3) {
    const bool hasm = val(v[C3_hasm]) != 0.0;
    nn[0] = v[C3_nn0]; nn[1] = v[C3_nnp]; nn[2] = hasm ? v[C3_nnm] : T(0.0);
    pc[0] = v[C3_Pc0]; pc[1] = v[C3_Pcp]; pc[2] = v[C3_Pcm];
  } else {
    nn[0] = v[C2_n1]; nn[1] = v[C2_np]; nn[2] = v[C2_nm];
    pc[0] = v[C2_Pc1]; pc[1] = v[C2_Pcp]; pc[2] = v[C2_Pcm];
  }
}

// dv = Xs * tau (the distance from the observer's argument U0)
template <class T, int CASE = -1>
KG_INLINE void radial_state(const RadialConsts<T>& rc, const JacobiLadder<T>& lad, const T& X, const T& dv,
                            bool do_phi_t, RState<T>& st) {
  const T* v = rc.v;
  const T& k = v[R_k];
  if (kfabs(val(dv)) <= 0.25 * val(v[R_K])) {
    // near the observer: addition theorem about u0 (DLMF 22.8.1-3), exact at tau = 0
    T s2, c2, d2;
    jacobi_eval(lad, dv, s2, c2, d2);
    const T& s1 = v[R_sn0];
    const T& c1 = v[R_cn0];
    const T& d1 = v[R_dn0];
    T inv = 1.0 / (1.0 - k * (s1 * s1) * (s2 * s2));
    st.s = (s1 * c2 * d2 - c1 * s2 * d1) * inv;
    st.c = (c1 * c2 + s1 * d1 * s2 * d2) * inv;
    st.d = (d1 * d2 + k * s1 * c1 * s2 * c2) * inv;
  } else {
    jacobi_eval(lad, X, st.s, st.c, st.d);
  }
  if (!do_phi_t) return;
  Amp<T> A = amp_from_jacobi(st.s, st.c, val(X), val(v[R_K]));
  LegendreFam<T, 3> fam;
  T nn[3], pc[3];
  radial_pi_params<T, CASE>(rc, nn, pc);
  legendre_family(A, k, v[R_K], v[R_EmK], nn, pc, true, fam);
  st.F = fam.F;
  st.EmF = fam.EmF;
  st.Pi[0] = fam.Pi[0];
  st.Pi[1] = fam.Pi[1];
  st.Pi[2] = fam.Pi[2];
  const int rcase = (CASE >= 0) ? CASE : rc.rcase;
  if (rcase == 3 || rcase == 4) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T a, b;
      if (rcase == 3) block_log_args<T, 3>(rc, i, st.s, st.d, a, b);
      else block_log_args<T, 4>(rc, i, st.s, st.d, a, b);
      st.lg[i] = log_ratio(a, b);
    }
  }
}

// r(tau) and, with do_phi_t, the building blocks I_+, I_-, I_1, I_2
// CASE = 2, 3, 4 fixes the radial case at compile time; CASE = -1 reads rc.rcase.
template <class T, int CASE = -1>
KG_INLINE RadialOut<T> radial_out(const RadialConsts<T>& rc, const T& tau, const RState<T>& st, bool do_phi_t) {
  RadialOut<T> o;
  const T* v = rc.v;
  const int rcase = (CASE >= 0) ? CASE : rc.rcase;
  const T& k = v[R_k];
  const T& sn = st.s;
  const T& cn = st.c;
  const T& dn = st.d;
  if (rcase == 4) {
    const T& a2 = v[C4_a2];
    const T& b1 = v[C4_b1];
    const T& g0 = v[C4_g0];
    // -a2 (g0 - sc)/(1 + g0 sc) - b1 with sc = sn/cn, written without 1/cn
    o.r = -a2 * (g0 * cn - sn) / (cn + g0 * sn) - b1;
    if (!do_phi_t) return o;
    const T& F = st.F;
    T E = F + st.EmF;
    T S1a0 = S1_from(g0, st.Pi[0], F, v[C4_q0], st.lg[0], v[C4_l0], v[C4_i0]);
    T S2a0 = S2_block(g0, S1a0, F, E, sn, cn, dn, k);
    T S1ap = S1_from(v[C4_gp], st.Pi[1], F, v[C4_qp], st.lg[1], v[C4_lp], v[C4_ip]);
    T S1am = S1_from(v[C4_gm], st.Pi[2], F, v[C4_qm], st.lg[2], v[C4_lm], v[C4_im]);
    T Pi1 = v[C4_P1f] * (S1a0 - v[C4_S1b0]);
    T Pi2 = v[C4_P2f] * (S2a0 - v[C4_S2b0]);
    T Pip = v[C4_Ppf] * (S1ap - v[C4_S1bp]);
    T Pim = v[C4_Pmf] * (S1am - v[C4_S1bm]);
    const T& aa1 = v[C4_aa1];
    o.I1 = aa1 * (-tau) - Pi1;
    o.I2 = (aa1 * aa1) * (-tau) - 2.0 * aa1 * Pi1 + Pi2;
    o.Ip = v[C4_Ipf] * (-tau - Pip);
    o.Im = v[C4_Imf] * (-tau - Pim);
  } else if (rcase == 3) {
    const T& AA = v[C3_AA];
    const T& BB = v[C3_BB];
    const T& rr1 = v[C3_rr1];
    const T& rr2 = v[C3_rr2];
    T num = (BB * rr2 - AA * rr1) + (BB * rr2 + AA * rr1) * cn;
    T den = (BB - AA) + (BB + AA) * cn;
    o.r = num / den;
    if (!do_phi_t) return o;
    const bool hasm = val(v[C3_hasm]) != 0.0;
    const T& F = st.F;
    T E = F + st.EmF;
    T R1a0 = R1_from(v[C3_al0], st.Pi[0], v[C3_p0], st.lg[0], v[C3_i0]);
    // the R2 block of R2_block with its per-pixel factors precomputed
    const T& al0 = v[C3_al0];
    T R2a0 = (F - v[C3_r2a] * (E - al0 * sn * dn / (1.0 + al0 * cn))) * v[C3_r2b] + v[C3_r2c] * R1a0;
    T R1ap = R1_from(v[C3_alp], st.Pi[1], v[C3_pp], st.lg[1], v[C3_ip]);
    T R1am = hasm ? R1_from(v[C3_alm], st.Pi[2], v[C3_pm], st.lg[2], v[C3_im]) : T(0.0);
    T Pi1 = v[C3_c0] * (R1a0 - v[C3_R1b0]);
    T Pi2 = (v[C3_c0] * v[C3_c0]) * (R2a0 - v[C3_R2b0]);
    T Pip = v[C3_cp] * (R1ap - v[C3_R1bp]);
    T Pim = v[C3_cm] * (R1am - v[C3_R1bm]);
    const T& pr = v[C3_pr];
    o.I1 = pr * (-tau) + Pi1;
    o.I2 = (pr * pr) * (-tau) + 2.0 * pr * Pi1 + v[C3_sqAB] * Pi2;
    o.Ip = -((BB + AA) * (-tau) + Pip) * v[C3_iIpd];
    o.Im = -((BB + AA) * (-tau) + Pim) * v[C3_iImd];
  } else {
    const T& rr3 = v[C2_rr3];
    const T& rr4 = v[C2_rr4];
    const T& rr31 = v[C2_rr31];
    const T& rr41 = v[C2_rr41];
    T sn2 = sn * sn;
    T den = rr31 - rr41 * sn2;
    o.r = (rr4 * rr31 - rr3 * rr41 * sn2) / den;
    if (!do_phi_t) return o;
    const T& rr43 = v[C2_rr43];
    T dsn2 = 2.0 * sn * cn * dn * v[C2_dX];
    // dr/dX / (r - r3) = -rr31 rr43 rr41 dsn2 / (den^2 (r - r3)), and
    // den (r - r3) = rr31 rr43, so H needs no further division
    T H = -(rr41 * dsn2) / den - v[C2_Ho];
    T Ea = st.F + st.EmF;
    T E = v[C2_sq] * (Ea - v[C2_Eb]);
    T Pi1 = v[C2_c] * (st.Pi[0] - v[C2_Pb1]);
    T Pip = v[C2_fp] * (st.Pi[1] - v[C2_Pbp]);
    T Pim = v[C2_fm] * (st.Pi[2] - v[C2_Pbm]);
    o.I1 = rr3 * (-tau) + rr43 * Pi1;
    o.I2 = H - v[C2_h12] * (-tau) - E;
    o.Ip = tau * v[C2_irp3] - Pip;
    o.Im = tau * v[C2_irm3] - Pim;
  }
  return o;
}

template <class T, int CASE = -1>
KG_INLINE RadialOut<T> radial_eval(const RadialConsts<T>& rc, const JacobiLadder<T>& lad, const T& tau,
                                   bool do_phi_t) {
  const T dv = rc.v[R_Xs] * tau;
  const T X = rc.v[R_U0] - dv;
  RState<T> st;
  radial_state<T, CASE>(rc, lad, X, dv, do_phi_t, st);
  return radial_out<T, CASE>(rc, tau, st, do_phi_t);
}

// ---------------------------------------------------------------------------
// total Mino time (kerr_raytracing_utils.py:382-548); assignments are made in
// the reference's order so overlapping masks resolve the same way.
// ---------------------------------------------------------------------------
template <class T> KG_INLINE T mino_total(const RayScalars<T>& c, const PixelSetup<T>& px) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const bool iv = !(val(px.ri[0]) == 0.0);
  const bool im3 = !(val(px.ri[2]) == 0.0);
  const T& rplus = c.rplus;
  const double r_o = c.r_o;
  T out = T(0.0);
  if (iv) {
    T a1 = px.ri[3], b1 = px.rr[3], a2 = px.ri[1], b2 = px.rr[1];
    T CC = ksqrt(sq(a1 - a2) + sq(b1 - b2));
    T DD = ksqrt(sq(a1 + a2) + sq(b1 - b2));
    T k4 = (4.0 * CC * DD) / sq(CC + DD);
    T g0 = ksqrt((4.0 * a2 * a2 - sq(CC - DD)) / (sq(CC + DD) - 4.0 * a2 * a2));
    T x4rp = (rplus + b1) / a2;
    T pref = 2.0 / (CC + DD);
    T K = ellip_K(k4);
    T sh, ch, so, co;
    atan_sum_sc(x4rp, g0, false, sh, ch);
    atan_sum_sc(c.r_o_inf ? T(0.0) : (r_o + b1) / a2, g0, c.r_o_inf, so, co);
    T Ihor = pref * legendre_F(legendre_base(amp_from_sc(sh, ch), k4), K);
    T Iobs = pref * legendre_F(legendre_base(amp_from_sc(so, co), k4), K);
    out = Iobs - Ihor;
  }
  if (im3 && !iv) {
    T a1 = px.ri[3], b1 = px.rr[3];
    T rr1 = px.rr[0], rr2 = px.rr[1];
    T rr21 = rr2 - rr1;
    T rrp1 = rplus - rr1, rrp2 = rplus - rr2;
    T AA = ksqrt(a1 * a1 + sq(b1 - rr2));
    T BB = ksqrt(a1 * a1 + sq(b1 - rr1));
    T k3 = (sq(AA + BB) - rr21 * rr21) / (4.0 * AA * BB);
    T x3rp = (AA * rrp1 - BB * rrp2) / (AA * rrp1 + BB * rrp2);
    T x3ro;
    if (c.r_o_inf) {
      x3ro = (AA - BB) / (AA + BB);
    } else {
      T rro1 = r_o - rr1, rro2 = r_o - rr2;
      x3ro = (AA * rro1 - BB * rro2) / (AA * rro1 + BB * rro2);
    }
    T pref = 1.0 / ksqrt(AA * BB);
    T K = ellip_K(k3);
    T Ihor = pref * legendre_F(legendre_base(amp_from_cos(x3rp), k3), K);
    T Iobs = pref * legendre_F(legendre_base(amp_from_sc(ksqrt((1.0 - x3ro) * (1.0 + x3ro)), x3ro), k3), K);
    out = Iobs - Ihor;
  }
  if (!im3) {
    const bool reg2 = val(px.rr[2]) < val(rplus);
    const bool reg1 = val(px.rr[2]) > val(rplus);
    if (reg2 || reg1) {
      T rr1 = px.rr[0], rr2 = px.rr[1], rr3 = px.rr[2], rr4 = px.rr[3];
      T rr31 = rr3 - rr1, rr32 = rr3 - rr2, rr41 = rr4 - rr1, rr42 = rr4 - rr2;
      T k2 = (rr32 * rr41) / (rr31 * rr42);
      T x2ro = c.r_o_inf ? ksqrt(rr31 / rr41) : ksqrt((rr31 * (r_o - rr4)) / (rr41 * (r_o - rr3)));
      T pref = 2.0 / ksqrt(rr31 * rr42);
      T K = ellip_K(k2);
      T Iobs = pref * legendre_F(legendre_base(amp_from_sc(x2ro, ksqrt((1.0 - x2ro) * (1.0 + x2ro))), k2), K);
      if (reg2) {
        T x2rp = ksqrt((rr31 * (rplus - rr4)) / (rr41 * (rplus - rr3)));
        T Ihor = pref * legendre_F(legendre_base(amp_from_sin(x2rp), k2), K);
        out = Iobs - Ihor;
      } else {
        out = 2.0 * Iobs;
      }
    }
  }
  return out;
}

// Full pixel setup: conserved quantities (with the spin-zero eta fix),
// angular turning points, roots, labels and tau_tot.  All per-pixel code
// is evaluated without FMA contraction so that its rounding is numpy's:
// u_+ = dth + sqrt(dth^2 + eta/a^2) cancels to ~1e-5 for near-polar rays,
// and the reference's own rounding there is what the samples inherit.
// TAU = false skips tau_tot (left 0), for callers that need the roots only.
template <class T, bool TAU = true>
KG_INLINE void pixel_setup(const RayScalars<T>& c, const T& alpha, const T& beta, PixelSetup<T>& px) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  px.status = 0;
  px.lam = -alpha * c.sin_th;
  px.eta = (alpha * alpha - c.a2) * c.cos2_th + beta * beta;
  if (c.minspin && val(px.eta) < 0.0) px.eta = T(kEP);
  px.s_o = (val(beta) >= 0.0) ? T(1.0) : (val(beta) < 0.0 ? T(-1.0) : T(0.0));
  uplus_uminus(c, px.lam, px.eta, px.up, px.um);
  const double ev = val(px.eta);
  px.thclass = ev > 0.0 ? 1 : (ev < 0.0 ? 2 : (ev == 0.0 ? 3 : 0));
  if (c.minspin && !(ev > 0.0)) px.status |= ST_MINSPIN_VORTICAL;
  radial_roots(c, px.lam, px.eta, px);
  px.tau_tot = TAU ? mino_total(c, px) : T(0.0);
}

// Mino-time grid value j of numpy.linspace(0, maxtaufrac*tau_tot, G)
// (kerr_raytracing_ana.py:94).  any_step_zero selects numpy's denormal-step
// formula, which then applies to every pixel.
KG_INLINE double mino_tau(int64_t j, int64_t G, double stop, bool any_step_zero) {
  if (G <= 1) return 0.0 * stop;
  if (j == G - 1) return stop;
  const double div = static_cast<double>(G - 1);
  if (any_step_zero) return (static_cast<double>(j) / div) * stop;
  return static_cast<double>(j) * (stop / div);
}

// Combination of kerr_raytracing_ana.py:186-188 and :627-632.
template <class T> struct CombineConsts {
  T fac_ph, cp, cm, fac_t, tp, tm, a2, lam;
};

template <class T> KG_INLINE CombineConsts<T> combine_consts(const RayScalars<T>& c, const T& lam) {
  CombineConsts<T> k;
  T dr = c.rplus - c.rminus;
  T hal = 0.5 * c.a * lam;
  k.fac_ph = (2.0 * c.a) / dr;
  k.cp = c.rplus - hal;
  k.cm = c.rminus - hal;
  k.fac_t = 4.0 / dr;
  k.tp = c.rplus * c.rplus - hal * c.rplus;
  k.tm = c.rminus * c.rminus - hal * c.rminus;
  k.a2 = c.a * c.a;
  k.lam = lam;
  return k;
}

template <class T> struct Sample {
  T t, r, th, ph, sig;
};

template <class T>
KG_INLINE void combine(const CombineConsts<T>& k, const T& tau, const RadialOut<T>& ro, const T& th, const T& Gph,
                       const T& Gt, Sample<T>& s) {
  // a = 0: the reference multiplies a finite bracket by exactly 0; near the
  // horizon-end sample our bracket can round past the pole of Pi_p to inf,
  // and 0 * inf would turn the reference's 0 into NaN
  T Iph = (val(k.fac_ph) == 0.0) ? k.fac_ph * 0.0 : k.fac_ph * (k.cp * ro.Ip - k.cm * ro.Im);
  T ItA = k.fac_t * (k.tp * ro.Ip - k.tm * ro.Im);
  T It = ItA + 4.0 * (-tau) + 2.0 * ro.I1 + ro.I2;
  s.sig = ro.I2 + k.a2 * Gt;
  s.t = It + k.a2 * Gt;
  s.ph = Iph + k.lam * Gph;
  s.r = ro.r;
  s.th = th;
}

// ---------------------------------------------------------------------------
// Pixel plan: everything the sample loop needs, as PLAN_N doubles per pixel
// (stored slot-major, [slot][npix], by the setup kernel).
// ---------------------------------------------------------------------------
enum PlanSlot : int {
  PL_TH = 0,
  PL_R = PL_TH + TH_NSLOT,
  PL_LAM = PL_R + R_NSLOT,
  PL_STOP,   // maxtaufrac * tau_tot
  PL_THK,    // theta kind
  PL_ETA0,   // eta == 0
  PL_RCASE,  // r_integrate case
  PL_N
};

template <class T> struct PixelConsts {
  ThetaConsts<T> th;
  RadialConsts<T> r;
  CombineConsts<T> cmb;
  JacobiLadder<T> lad_th, lad_r;
  T stop;
};

// The plan in two parts: theta, lambda and the grid stop; the radial
// constants (PL_R .. PL_R + R_NSLOT - 1 and PL_RCASE).
KG_INLINE void plan_pack_theta(const ThetaConsts<double>& tc, double lam, double stop, double* plan) {
  for (int i = 0; i < TH_NSLOT; ++i) plan[PL_TH + i] = tc.v[i];
  plan[PL_LAM] = lam;
  plan[PL_STOP] = stop;
  plan[PL_THK] = static_cast<double>(tc.kind);
  plan[PL_ETA0] = static_cast<double>(tc.eta_zero);
}

KG_INLINE void plan_pack_radial(const RadialConsts<double>& rc, double* plan) {
  for (int i = 0; i < R_NSLOT; ++i) plan[PL_R + i] = rc.v[i];
  plan[PL_RCASE] = static_cast<double>(rc.rcase);
}

KG_INLINE bool plan_slot_radial(int s) { return (s >= PL_R && s < PL_R + R_NSLOT) || s == PL_RCASE; }

KG_INLINE void plan_pack(const ThetaConsts<double>& tc, const RadialConsts<double>& rc, double lam, double stop,
                         double* plan) {
  plan_pack_theta(tc, lam, stop, plan);
  plan_pack_radial(rc, plan);
}

KG_INLINE void plan_unpack(const RayScalars<double>& c, const double* plan, PixelConsts<double>& pc) {
  for (int i = 0; i < TH_NSLOT; ++i) pc.th.v[i] = plan[PL_TH + i];
  for (int i = 0; i < R_NSLOT; ++i) pc.r.v[i] = plan[PL_R + i];
  pc.th.kind = static_cast<int>(plan[PL_THK]);
  pc.th.eta_zero = static_cast<int>(plan[PL_ETA0]);
  pc.r.rcase = static_cast<int>(plan[PL_RCASE]);
  pc.stop = plan[PL_STOP];
  pc.cmb = combine_consts(c, plan[PL_LAM]);
  jacobi_ladder(pc.th.v[TH_mp], pc.lad_th);
  jacobi_ladder(pc.r.v[R_k], pc.lad_r);
}

// One sample of kerr_raytracing_ana.coords_at_tau for one pixel.
// theta_on = false reproduces the reference's zero theta outputs (e.g. for
// eta == 0 when no pixel of the image has eta > 0, kerr_raytracing_ana.py:217).
template <class T>
KG_INLINE Sample<T> eval_sample(const PixelConsts<T>& pc, const T& tau, bool do_phi_t, bool theta_on) {
  T th = T(0.0), Gph = T(0.0), Gt = T(0.0);
  if (theta_on && pc.th.kind != THK_NONE) theta_eval(pc.th, pc.lad_th, tau, do_phi_t, th, Gph, Gt);
  RadialOut<T> ro = radial_eval(pc.r, pc.lad_r, tau, do_phi_t);
  Sample<T> s;
  if (do_phi_t) {
    combine(pc.cmb, tau, ro, th, Gph, Gt, s);
  } else {
    s.r = ro.r;
    s.th = th;
    s.t = s.ph = s.sig = T(0.0);
  }
  return s;
}

// Setup of one pixel into its plan: returns the PixelSetup (labels, status).
KG_INLINE PixelSetup<double> plan_pixel(const RayScalars<double>& c, double alpha, double beta, bool do_phi_t,
                                        double* plan) {
  PixelSetup<double> px;
  pixel_setup(c, alpha, beta, px);
  ThetaConsts<double> tc;
  RadialConsts<double> rc;
  theta_consts(c, px, do_phi_t, tc);
  radial_consts(c, px, do_phi_t, rc);
  plan_pack(tc, rc, px.lam, c.maxtaufrac * px.tau_tot, plan);
  return px;
}

// The plan without its radial constants (plan_pack_theta's slots).
KG_INLINE PixelSetup<double> plan_pixel_theta(const RayScalars<double>& c, double alpha, double beta, bool do_phi_t,
                                              double* plan) {
  PixelSetup<double> px;
  pixel_setup(c, alpha, beta, px);
  ThetaConsts<double> tc;
  theta_consts(c, px, do_phi_t, tc);
  plan_pack_theta(tc, px.lam, c.maxtaufrac * px.tau_tot, plan);
  return px;
}

// The radial constants alone (plan_pack_radial's slots), from the same
// per-pixel set-up, so that they equal plan_pixel's bit for bit.
KG_INLINE void plan_pixel_radial(const RayScalars<double>& c, double alpha, double beta, bool do_phi_t, double* plan) {
  PixelSetup<double> px;
  pixel_setup<double, false>(c, alpha, beta, px);
  RadialConsts<double> rc;
  radial_consts(c, px, do_phi_t, rc);
  plan_pack_radial(rc, plan);
}

// Geodesics.nmax_eq (kerr_raytracing_utils.py:63-67, 605-684) at the last
// Mino-time sample tau_last; uses eta without the spin-zero fix.
KG_INLINE int nmax_equatorial_at(const RayScalars<double>& c, double alpha, double beta, double tau_last) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double lam = -alpha * c.sin_th;
  double eta = (alpha * alpha - c.a2) * c.cos2_th + beta * beta;
  if (!(eta > 0.0)) return (eta <= 0.0) ? -2 : 0;
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
  double K = ellip_K(ratio);
  double F0 = legendre_F(legendre_base(amp_from_sin(c.cos_th / ksqrt(up)), ratio), K);
  double w = ksqrt(-a2um);
  double x = tau_last * w;
  double n;
  if (c.h > 0.0) {
    n = (beta <= 0.0) ? kfloor((x - F0) / (2.0 * K)) : kfloor((x + F0) / (2.0 * K)) - 1.0;
  } else {
    n = (beta >= 0.0) ? kfloor((x + F0) / (2.0 * K)) : kfloor((x - F0) / (2.0 * K)) - 1.0;
  }
  return static_cast<int>(n);
}

}  // namespace kg
