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
inline CD frcp(const CD& x) { ++g_ops[DIV]; return CD(1.0 / x.v); }
inline CD kpow(const CD& x, double y) { ++g_ops[POW]; return CD(std::pow(x.v, y)); }

}  // namespace kgcount

#include "kg_geodesic.hpp"
#include "kg_host_passes.hpp"
#include "kg_sample.hpp"

using kgcount::CD;
using namespace kg;


extern "C" {

int kgc_num_ops(void) { return kgcount::NOPS; }

// Per-sample op counts summed over npix pixels x ngeo linspace samples, as
// the sample kernels execute them: the anchor / walk / fixup passes
// (kg_host_passes.hpp: direct evaluations at the step, the segment starts,
// the last sample and from each refused step on; the walk elsewhere) with
// ngeo split into `chunks` ranges per pixel as the kernels' grid does.
// Separately, the per-pixel setup counts.  sc as in kg_hostcheck.cpp.
void kgc_count(const double* sc, const double* alpha, const double* beta, int64_t npix, int64_t ngeo, int chunks,
               uint64_t* per_sample_ops, uint64_t* setup_ops) {
  RayScalars<double> c = make_scalars(sc[0], sc[2], sc[1], sc[3]);
  c.a2 = sc[4]; c.sin_th = sc[5]; c.cos_th = sc[6]; c.cos2_th = sc[7]; c.rplus = sc[8]; c.rminus = sc[9];
  RayScalars<CD> cc;
  cc.a = CD(c.a); cc.a2 = CD(c.a2); cc.sin_th = CD(c.sin_th); cc.cos_th = CD(c.cos_th);
  cc.cos2_th = CD(c.cos2_th); cc.rplus = CD(c.rplus); cc.rminus = CD(c.rminus);
  cc.r_o = c.r_o; cc.maxtaufrac = c.maxtaufrac; cc.h = c.h; cc.minspin = c.minspin;
  cc.spin_gt_min = c.spin_gt_min; cc.r_o_inf = c.r_o_inf;
  std::vector<double> plan(static_cast<size_t>(npix) * PL_N);
  uint64_t setup_acc[kgcount::NOPS] = {0}, samp_acc[kgcount::NOPS] = {0};
  uint32_t gflags = 0;
  for (int64_t i = 0; i < npix; ++i) {
    double pl[PL_N];
    PixelSetup<double> px = plan_pixel(c, alpha[i], beta[i], true, pl);
    for (int s = 0; s < PL_N; ++s) plan[static_cast<size_t>(s) * npix + i] = pl[s];
    if (px.eta > 0.0) gflags |= GF_ANY_ETA_POS;
    if (ngeo > 1 && pl[PL_STOP] / static_cast<double>(ngeo - 1) == 0.0) gflags |= GF_ANY_STEP_ZERO;
    // setup cost: theta/radial constants and ladders, counted on the CD type
    std::memset(kgcount::g_ops, 0, sizeof(kgcount::g_ops));
    PixelSetup<CD> pxc;
    pixel_setup(cc, CD(alpha[i]), CD(beta[i]), pxc);
    ThetaConsts<CD> tcc;
    RadialConsts<CD> rcc;
    theta_consts(cc, pxc, true, tcc);
    radial_consts(cc, pxc, true, rcc);
    for (int q = 0; q < kgcount::NOPS; ++q) setup_acc[q] += kgcount::g_ops[q];
  }
  // per-sample cost: the kernels' loops on the counting type
  std::vector<double> t(static_cast<size_t>(ngeo) * npix), r(t.size()), th(t.size()), ph(t.size());
  SampleOut o{nullptr, nullptr, t.data(), r.data(), th.data(), ph.data(), nullptr, npix, ngeo};
  if (chunks < 1) chunks = 1;
  std::memset(kgcount::g_ops, 0, sizeof(kgcount::g_ops));
  std::vector<uint32_t> st(static_cast<size_t>(npix), 0u);
  int64_t nfix[3];
  three_passes<true, CD>(cc, plan.data(), gflags, o, chunks, st.data(), nfix);
  for (int q = 0; q < kgcount::NOPS; ++q) samp_acc[q] += kgcount::g_ops[q];
  for (int q = 0; q < kgcount::NOPS; ++q) {
    per_sample_ops[q] = samp_acc[q];
    setup_ops[q] = setup_acc[q];
  }
}

}  // extern "C"
