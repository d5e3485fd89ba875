// Algorithmic operation counter for the raytracer's per-sample work.
//
// Instantiates the device source (kg_geodesic.hpp) on the CPU with a
// counting scalar type and reports, per sample, how many FP64 additions,
// multiplications, divisions, square roots and calls of each transcendental
// the algorithm performs.  bench.py turns the counts into the algorithmic
// FLOPs per sample used for the roofline (weights documented there).
// Measurement instrument only; never loaded by the product modules.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "kg_math.hpp"

namespace kgcount {

enum Op { ADD, MUL, DIV, SQRT, SIN, COS, TAN, ASIN, ACOS, ATAN, ATAN2, LOG, LOG1P, EXP, CBRT, POW, NOPS };
static uint64_t g_ops[NOPS];

struct CD {
  double v;
  CD() : v(0.0) {}
  CD(double x) : v(x) {}
};

inline CD operator+(const CD& a, const CD& b) { ++g_ops[ADD]; return CD(a.v + b.v); }
inline CD operator-(const CD& a, const CD& b) { ++g_ops[ADD]; return CD(a.v - b.v); }
inline CD operator*(const CD& a, const CD& b) { ++g_ops[MUL]; return CD(a.v * b.v); }
inline CD operator/(const CD& a, const CD& b) { ++g_ops[DIV]; return CD(a.v / b.v); }
inline CD operator-(const CD& a) { return CD(-a.v); }
inline CD operator+(double a, const CD& b) { return CD(a) + b; }
inline CD operator-(double a, const CD& b) { return CD(a) - b; }
inline CD operator*(double a, const CD& b) { return CD(a) * b; }
inline CD operator/(double a, const CD& b) { return CD(a) / b; }
inline CD operator+(const CD& a, double b) { return a + CD(b); }
inline CD operator-(const CD& a, double b) { return a - CD(b); }
inline CD operator*(const CD& a, double b) { return a * CD(b); }
inline CD operator/(const CD& a, double b) { return a / CD(b); }

inline double val(const CD& x) { return x.v; }
inline CD ksqrt(const CD& x) { ++g_ops[SQRT]; return CD(std::sqrt(x.v)); }
inline CD kcbrt(const CD& x) { ++g_ops[CBRT]; return CD(std::cbrt(x.v)); }
inline CD ksin(const CD& x) { ++g_ops[SIN]; return CD(std::sin(x.v)); }
inline CD kcos(const CD& x) { ++g_ops[COS]; return CD(std::cos(x.v)); }
inline CD ktan(const CD& x) { ++g_ops[TAN]; return CD(std::tan(x.v)); }
inline CD kasin(const CD& x) { ++g_ops[ASIN]; return CD(std::asin(x.v)); }
inline CD kacos(const CD& x) { ++g_ops[ACOS]; return CD(std::acos(x.v)); }
inline CD katan(const CD& x) { ++g_ops[ATAN]; return CD(std::atan(x.v)); }
inline CD katan2(const CD& y, const CD& x) { ++g_ops[ATAN2]; return CD(std::atan2(y.v, x.v)); }
inline CD klog(const CD& x) { ++g_ops[LOG]; return CD(std::log(x.v)); }
inline CD klog1p(const CD& x) { ++g_ops[LOG1P]; return CD(std::log1p(x.v)); }
inline CD kexp(const CD& x) { ++g_ops[EXP]; return CD(std::exp(x.v)); }
inline CD kfabs(const CD& x) { return CD(std::fabs(x.v)); }
inline CD kpow(const CD& x, double y) { ++g_ops[POW]; return CD(std::pow(x.v, y)); }

}  // namespace kgcount

#include "kg_geodesic.hpp"

using kgcount::CD;
using namespace kg;

namespace {

template <class A, class B> void copy_consts(const PixelConsts<A>& s, PixelConsts<B>& d) {
  for (int i = 0; i < TH_NSLOT; ++i) d.th.v[i] = B(s.th.v[i]);
  for (int i = 0; i < R_NSLOT; ++i) d.r.v[i] = B(s.r.v[i]);
  d.th.kind = s.th.kind;
  d.th.eta_zero = s.th.eta_zero;
  d.r.rcase = s.r.rcase;
  d.stop = B(s.stop);
  d.cmb.fac_ph = B(s.cmb.fac_ph); d.cmb.cp = B(s.cmb.cp); d.cmb.cm = B(s.cmb.cm);
  d.cmb.fac_t = B(s.cmb.fac_t); d.cmb.tp = B(s.cmb.tp); d.cmb.tm = B(s.cmb.tm);
  d.cmb.a2 = B(s.cmb.a2); d.cmb.lam = B(s.cmb.lam);
  const JacobiLadder<A>* ls[2] = {&s.lad_th, &s.lad_r};
  JacobiLadder<B>* ld[2] = {&d.lad_th, &d.lad_r};
  for (int q = 0; q < 2; ++q) {
    for (int i = 0; i < kMaxLadder; ++i) ld[q]->k[i] = B(ls[q]->k[i]);
    ld[q]->aN = B(ls[q]->aN);
    ld[q]->K = B(ls[q]->K);
    ld[q]->n = ls[q]->n;
    ld[q]->capped = ls[q]->capped;
  }
}

}  // namespace

extern "C" {

int kgc_num_ops(void) { return kgcount::NOPS; }

// Per-sample op counts summed over npix pixels x ngeo linspace samples
// (plus, separately, the per-pixel setup counts).  sc as in kg_hostcheck.cpp.
void kgc_count(const double* sc, const double* alpha, const double* beta, int64_t npix, int64_t ngeo,
               uint64_t* per_sample_ops, uint64_t* setup_ops) {
  RayScalars<double> c = make_scalars(sc[0], sc[2], sc[1], sc[3]);
  c.a2 = sc[4]; c.sin_th = sc[5]; c.cos_th = sc[6]; c.cos2_th = sc[7]; c.rplus = sc[8]; c.rminus = sc[9];
  std::vector<double> plan(PL_N);
  uint64_t setup_acc[kgcount::NOPS] = {0}, samp_acc[kgcount::NOPS] = {0};
  for (int64_t i = 0; i < npix; ++i) {
    PixelSetup<double> px = plan_pixel(c, alpha[i], beta[i], true, plan.data());
    (void)px;
    // setup cost: theta/radial constants and ladders, counted on the CD type
    std::memset(kgcount::g_ops, 0, sizeof(kgcount::g_ops));
    RayScalars<CD> cc;
    cc.a = CD(c.a); cc.a2 = CD(c.a2); cc.sin_th = CD(c.sin_th); cc.cos_th = CD(c.cos_th);
    cc.cos2_th = CD(c.cos2_th); cc.rplus = CD(c.rplus); cc.rminus = CD(c.rminus);
    cc.r_o = c.r_o; cc.maxtaufrac = c.maxtaufrac; cc.h = c.h; cc.minspin = c.minspin;
    cc.spin_gt_min = c.spin_gt_min; cc.r_o_inf = c.r_o_inf;
    PixelSetup<CD> pxc;
    pixel_setup(cc, CD(alpha[i]), CD(beta[i]), pxc);
    ThetaConsts<CD> tcc;
    RadialConsts<CD> rcc;
    theta_consts(cc, pxc, true, tcc);
    radial_consts(cc, pxc, true, rcc);
    JacobiLadder<CD> l1, l2;
    jacobi_ladder(tcc.v[TH_mp], l1);
    jacobi_ladder(rcc.v[R_k], l2);
    for (int q = 0; q < kgcount::NOPS; ++q) setup_acc[q] += kgcount::g_ops[q];
    // per-sample cost
    PixelConsts<double> pc;
    plan_unpack(c, plan.data(), pc);
    PixelConsts<CD> pcc;
    copy_consts(pc, pcc);
    std::memset(kgcount::g_ops, 0, sizeof(kgcount::g_ops));
    for (int64_t j = 0; j < ngeo; ++j) {
      double tj = mino_tau(j, ngeo, pc.stop, false);
      Sample<CD> s = eval_sample(pcc, CD(tj), true, true);
      (void)s;
    }
    for (int q = 0; q < kgcount::NOPS; ++q) samp_acc[q] += kgcount::g_ops[q];
  }
  for (int q = 0; q < kgcount::NOPS; ++q) {
    per_sample_ops[q] = samp_acc[q];
    setup_ops[q] = setup_acc[q];
  }
}

}  // extern "C"
