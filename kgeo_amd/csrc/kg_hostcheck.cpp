// CPU instantiation of the device math (kg_math.hpp / kg_geodesic.hpp).
//
// Test and measurement instrument only: it lets the CPU test suite check
// the exact source the HIP kernels compile (Carlson, Legendre, Jacobi,
// roots, per-sample evaluation) against the oracle without a GPU, and it
// backs the algorithmic FLOP counter.  The product modules never load it:
// kgeo_amd's raytrace entry points require the HIP library and raise if it
// is missing.

#include <cmath>
#include <cstdint>
#include <climits>
#include <cstring>
#include <vector>

#include "kg_counts.hpp"
#include "kg_dual.hpp"
#include "kg_emission.hpp"
#include "kg_equatorial.hpp"
#include "kg_geodesic.hpp"
#include "kg_host_passes.hpp"
#include "kg_lensing.hpp"
#include "kg_sample.hpp"

using namespace kg;

namespace {

RayScalars<double> scalars_from(const double* s) {
  // s: a, r_o, th_o, maxtaufrac, a2, sin_th, cos_th, cos2_th, rplus, rminus
  RayScalars<double> c = make_scalars(s[0], s[2], s[1], s[3]);
  c.a2 = s[4];
  c.sin_th = s[5];
  c.cos_th = s[6];
  c.cos2_th = s[7];
  c.rplus = s[8];
  c.rminus = s[9];
  return c;
}

}  // namespace


extern "C" {

void kgh_rf(const double* x, const double* y, const double* z, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = carlson_rf(x[i], y[i], z[i]);
}

void kgh_rd(const double* x, const double* y, const double* z, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = carlson_rd(x[i], y[i], z[i]);
}

void kgh_rj(const double* x, const double* y, const double* z, const double* p, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double rf = carlson_rf(x[i], y[i], z[i]);
    out[i] = carlson_rj_any(x[i], y[i], z[i], p[i], rf);
  }
}

void kgh_rc(const double* x, const double* y, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = carlson_rc(x[i], y[i]);
}

void kgh_ellipj(const double* u, const double* m, int64_t n, double* sn, double* cn, double* dn, double* K) {
  for (int64_t i = 0; i < n; ++i) {
    JacobiLadder<double> L;
    jacobi_ladder(m[i], L);
    jacobi_eval(L, u[i], sn[i], cn[i], dn[i]);
    K[i] = L.K;
  }
}

// Legendre F, E, Pi at arbitrary amplitude phi (reference conventions,
// including ellip_pi_arr's special cases).
void kgh_legendre(const double* nn, const double* phi, const double* m, int64_t n, double* F, double* E, double* P) {
  for (int64_t i = 0; i < n; ++i) {
    Amp<double> A = amp_from_angle(phi[i]);
    LegendreBase<double> L = legendre_base(A, m[i]);
    double K = ellip_K(m[i]);
    double EmK = ellip_EmK(m[i]);
    F[i] = legendre_F(L, K);
    E[i] = F[i] + legendre_EmF(L, m[i], EmK);
    double Pc = ellip_pi_complete(nn[i], m[i], K);
    P[i] = ellip_pi_amp(L, nn[i], m[i], Pc);
  }
}

// Setup: labels, tau_tot, roots, u_+-.
void kgh_setup(const double* sc, const double* alpha, const double* beta, int64_t n, int32_t* rclass, int32_t* thclass,
               int32_t* rcase, uint32_t* status, double* tau_tot, double* roots8, double* upum) {
  RayScalars<double> c = scalars_from(sc);
  for (int64_t i = 0; i < n; ++i) {
    PixelSetup<double> px;
    pixel_setup(c, alpha[i], beta[i], px);
    rclass[i] = px.rclass;
    thclass[i] = px.thclass;
    rcase[i] = px.rcase;
    status[i] = px.status;
    tau_tot[i] = px.tau_tot;
    for (int k = 0; k < 4; ++k) {
      roots8[8 * i + 2 * k] = px.rr[k];
      roots8[8 * i + 2 * k + 1] = px.ri[k];
    }
    upum[2 * i] = px.up;
    upum[2 * i + 1] = px.um;
  }
}


// Full raytrace of the kernels' algorithm on the CPU: the same plan layout
// ([PL_N][npix] slot-major), image-wide flags and per-pixel sample code
// (kg_sample.hpp) as kg_setup and the sample kernels, with ngeo split into
// `chunks` contiguous ranges per pixel as the kernels' grid does.
//   passes = 3: the anchor / walk / fixup passes (the kernels' path for the
//               linspace grid with ngeo > 2); nfix[3] (optional) receives the
//               theta and radial fixup-queue lengths and the radial repair
//               items;
//   passes = 1: one combined loop per coordinate (theta_pixel, radial_pixel:
//               the kernels' path for caller-given Mino times).
// Outputs [ngeo][npix]; tau_in (optional, [ngeo][npix]) replaces the linspace
// grid, which also makes every sample a direct evaluation.  status (optional)
// receives the kernels' per-pixel bits.
void kgh_raytrace_passes(const double* sc, const double* alpha, const double* beta, int64_t npix, int64_t ngeo,
                         int do_phi_t, const double* tau_in, int chunks, int passes, double* tau, double* t, double* r,
                         double* th, double* ph, double* sig, int32_t* nmax_eq, uint32_t* status, int64_t* nfix) {
  RayScalars<double> c = scalars_from(sc);
  std::vector<double> plan(static_cast<size_t>(npix) * PL_N);
  std::vector<int8_t> rcl(npix), tcl(npix);
  std::vector<double> tt(npix);
  std::vector<uint32_t> st(npix, 0u);
  uint32_t gflags = 0;
  // the plan in the two halves the plan kernels write (kg_setup_theta,
  // kg_setup_radial)
  for (int64_t i = 0; i < npix; ++i) {
    SetupOut so = setup_pixel<true>(c, ngeo, do_phi_t != 0, alpha[i], beta[i], plan.data(), npix, i, rcl.data(),
                                    tcl.data(), nmax_eq, tt.data(), st.data(), i);
    setup_radial_pixel(c, do_phi_t != 0, alpha[i], beta[i], plan.data(), npix, i);
    if (so.eta_pos) gflags |= GF_ANY_ETA_POS;
    if (so.step_zero) gflags |= GF_ANY_STEP_ZERO;
  }
  const int64_t G = ngeo;
  if (chunks < 1) chunks = 1;
  SampleOut o{tau_in, tau, t, r, th, ph, sig, npix, G};
  std::vector<double> tbuf, pbuf;
  if (!do_phi_t) {
    o.t = nullptr;
    o.ph = nullptr;
  }
  if (passes == 3 && !tau_in && G > 2) {
    int64_t nf[3] = {0, 0, 0};
    if (do_phi_t) three_passes<true, double>(c, plan.data(), gflags, o, chunks, st.data(), nf);
    else three_passes<false, double>(c, plan.data(), gflags, o, chunks, st.data(), nf);
    if (nfix) {
      nfix[0] = nf[0];
      nfix[1] = nf[1];
      nfix[2] = nf[2];
    }
    if (status)
      for (int64_t i = 0; i < npix; ++i) status[i] = st[i];
    return;
  }
  for (int64_t i = 0; i < npix; ++i) {
    for (int ch = 0; ch < chunks; ++ch) {
      const int64_t j0 = chunk_start(G, ch, chunks), j1 = chunk_start(G, ch + 1, chunks);
      uint32_t bits;
      if (do_phi_t) {
        bits = theta_pixel<true>(plan.data(), npix, i, gflags, o, i, j0, j1);
        bits |= radial_pixel<true>(c, plan.data(), npix, i, gflags, o, i, j0, j1);
      } else {
        bits = theta_pixel<false>(plan.data(), npix, i, gflags, o, i, j0, j1);
        bits |= radial_pixel<false>(c, plan.data(), npix, i, gflags, o, i, j0, j1);
      }
      st[i] |= bits;
    }
  }
  if (status)
    for (int64_t i = 0; i < npix; ++i) status[i] = st[i];
}

void kgh_raytrace_loops(const double* sc, const double* alpha, const double* beta, int64_t npix, int64_t ngeo,
                        int do_phi_t, const double* tau_in, int chunks, double* tau, double* t, double* r, double* th,
                        double* ph, double* sig, int32_t* nmax_eq, uint32_t* status) {
  kgh_raytrace_passes(sc, alpha, beta, npix, ngeo, do_phi_t, tau_in, chunks, 3, tau, t, r, th, ph, sig, nmax_eq,
                      status, nullptr);
}

// Backwards-compatible entry: all samples of a pixel in one range.
void kgh_raytrace(const double* sc, const double* alpha, const double* beta, int64_t npix, int64_t ngeo, int do_phi_t,
                  const double* tau_in, double* tau, double* t, double* r, double* th, double* ph, double* sig,
                  int32_t* nmax_eq) {
  kgh_raytrace_loops(sc, alpha, beta, npix, ngeo, do_phi_t, tau_in, 1, tau, t, r, th, ph, sig, nmax_eq, nullptr);
}

// Count formulas of the kg_counts / kg_turnings kernels (kg_counts.hpp) at
// Mino times tau [ntau][n]; every output [ntau][n] (any may be NULL).
void kgh_counts(const double* sc, const double* alpha, const double* beta, int64_t n, const double* tau,
                int64_t ntau, double* n_eq, double* n_pol, double* n_ang, int32_t* turns_th, int8_t* turns_r) {
  RayScalars<double> c = scalars_from(sc);
  const bool north = sc[2] < kHalfPi;
  for (int64_t i = 0; i < n; ++i) {
    const CountConsts k = count_consts(c, alpha[i], beta[i]);
    const RadialTurn rt = radial_turn_consts(c, alpha[i], beta[i]);
    for (int64_t j = 0; j < ntau; ++j) {
      const int64_t o = j * n + i;
      const double tj = tau[o];
      if (n_eq) n_eq[o] = n_equatorial_at(k, north, beta[i], tj);
      if (n_pol) n_pol[o] = n_poloidal_at(k, tj);
      const double na = n_angular_at(k, beta[i], tj);
      if (n_ang) n_ang[o] = na;
      if (turns_th) turns_th[o] = (na == na && std::fabs(na) < 2.0e9) ? static_cast<int32_t>(std::floor(na)) : INT32_MIN;
      if (turns_r) turns_r[o] = static_cast<int8_t>(radial_turns_at(rt, tj));
    }
  }
}

// rho_of_req (kg_lensing.hpp) on the CPU for one curve: rho[i], status[i]
// at varphi[i]; sc as for the other entries (r_o = inf for the reference's).
void kgh_rho_of_req(const double* sc, double req, int mbar, const double* varphi, int64_t n, double* rho,
                    int32_t* status) {
  RayScalars<double> c = scalars_from(sc);
  for (int64_t i = 0; i < n; ++i) status[i] = rho_of_req_at(c, req, mbar, varphi[i], &rho[i]);
}

// r_equatorial2 (kg_lensing.hpp) on the CPU at screen points (rho[i], varphi[i]).
void kgh_r_equatorial2(const double* sc, int mbar, const double* rho, const double* varphi, int64_t n, double* r) {
  RayScalars<double> c = scalars_from(sc);
  for (int64_t i = 0; i < n; ++i) r[i] = r_equatorial2(c, rho[i], std::cos(varphi[i]), std::sin(varphi[i]), mbar);
}

// Equatorial image (kg_equatorial.hpp) on the CPU: per-subring arrays are
// [mbar_max+1][n].
void kgh_equatorial(const double* sc, const double* alpha, const double* beta, int64_t n, int mbar_max, double p,
                    double* image, double* I, double* g, double* rs, double* Ir, double* Imax, int32_t* Nmax,
                    int32_t* zmask) {
  RayScalars<double> c = scalars_from(sc);
  const double ri = r_isco(c.a);
  for (int64_t i = 0; i < n; ++i) {
    EqPixel<double> e;
    equatorial_pixel(c, alpha[i], beta[i], p, ri, mbar_max, e);
    image[i] = e.image;
    for (int m = 0; m <= mbar_max; ++m) {
      I[m * n + i] = e.I[m];
      g[m * n + i] = e.g[m];
      rs[m * n + i] = e.r_s[m];
      Ir[m * n + i] = e.Ir[m];
    }
    Imax[i] = e.Imax;
    Nmax[i] = e.Nmax;
    zmask[i] = e.zero_mask;
  }
}

// dImage/d(a, th_o, p) per pixel with forward-mode duals: jac [3][n].
void kgh_equatorial_jac(double a, double th_o, double r_o, double p, const double* alpha, const double* beta,
                        int64_t n, int mbar_max, double* image, double* jac) {
  using D = kgd::Dual<3>;
  const D ad = D::seed(a, 0), thd = D::seed(th_o, 1), pd = D::seed(p, 2);
  RayScalars<D> c = make_scalars(ad, thd, r_o, 1.0);
  const D ri = r_isco(ad);
  for (int64_t i = 0; i < n; ++i) {
    EqPixel<D> e;
    equatorial_pixel(c, D(alpha[i]), D(beta[i]), pd, ri, mbar_max, e);
    image[i] = e.image.v;
    for (int k = 0; k < 3; ++k) jac[k * n + i] = e.image.d[k];
  }
}

// Polarized image (kg_emission.hpp) on the CPU, outputs [n][nsub] as
// kgeo_equatorial_stokes writes them.
void kgh_stokes(const double* sc, const kgeo_emission_model* M, const double* alpha, const double* beta, int64_t n,
                int nsub, double* I, double* Q, double* U, double* g, double* rs, double* sinthb, double* Ir,
                double* nmax, double* npol, double* lp, double* Ie) {
  RayScalars<double> c = scalars_from(sc);
  const bool north = sc[2] < kHalfPiD;
  const double bsq0 =
      (M->emissivity.kind == KGEO_EMIS_THERMAL) ? field_bsq(M->bfield, M->velocity, c.a, M->emissivity.Rb, equator()) : 0.0;
  StokesArrays o{I, Q, U, g, rs, sinthb, Ir, nmax, npol, lp, Ie};
  for (int64_t i = 0; i < n; ++i) stokes_pixel(c, north, *M, alpha[i], beta[i], nsub, bsq0, i, o);
}

// The emission models at points (kgeo_model_eval on the CPU); th NULL is the
// equator.
void kgh_model_eval(const kgeo_emission_model* M, double a, const double* r, const double* th, int64_t n,
                    const double* g, const double* sinthb, const double* bmag, double* u, double* tet, double* B,
                    double* omega, double* bsq, double* j) {
  for (int64_t i = 0; i < n; ++i) {
    const Angle an = th ? angle_of(th[i]) : equator();
    model_point(*M, a, r[i], an, i, n, g, sinthb, bmag, u, tet, B, omega, bsq, j);
  }
}

// The plan of npix pixels written whole (kg_setup) and in its two halves
// (kg_setup_theta, kg_setup_radial): the number of slots that differ in any
// bit, and of label, tau_tot, nmax_eq or status values that differ.
int64_t kgh_plan_split_mismatches(const double* sc, const double* alpha, const double* beta, int64_t npix,
                                  int64_t ngeo) {
  RayScalars<double> c = scalars_from(sc);
  std::vector<double> p1(static_cast<size_t>(npix) * PL_N, -1.0), p2(static_cast<size_t>(npix) * PL_N, -2.0);
  std::vector<int8_t> r1(npix), r2(npix), t1(npix), t2(npix);
  std::vector<double> u1(npix), u2(npix);
  std::vector<int32_t> n1(npix), n2(npix);
  std::vector<uint32_t> s1(npix, 0u), s2(npix, 0u);
  int64_t bad = 0;
  for (int64_t i = 0; i < npix; ++i) {
    SetupOut a = setup_pixel(c, ngeo, true, alpha[i], beta[i], p1.data(), npix, i, r1.data(), t1.data(), n1.data(),
                             u1.data(), s1.data(), i);
    SetupOut b = setup_pixel<true>(c, ngeo, true, alpha[i], beta[i], p2.data(), npix, i, r2.data(), t2.data(),
                                   n2.data(), u2.data(), s2.data(), i);
    setup_radial_pixel(c, true, alpha[i], beta[i], p2.data(), npix, i);
    bad += (a.eta_pos != b.eta_pos) + (a.step_zero != b.step_zero);
    bad += (r1[i] != r2[i]) + (t1[i] != t2[i]) + (n1[i] != n2[i]) + (s1[i] != s2[i]);
    bad += std::memcmp(&u1[i], &u2[i], sizeof(double)) != 0;
  }
  for (size_t k = 0; k < p1.size(); ++k) bad += std::memcmp(&p1[k], &p2[k], sizeof(double)) != 0;
  return bad;
}

// Special functions: which 0 = Li2(x) (x <= 1), 1 = Lambert W0(x), 2 = BZ
// f(r) with f'(r) in out2 (the field models), 3 = the theta outputs' arccos.
void kgh_special(int which, const double* x, int64_t n, double* out, double* out2) {
  for (int64_t i = 0; i < n; ++i) {
    if (which == 0) out[i] = li2(x[i]);
    else if (which == 1) out[i] = lambert_w0(x[i]);
    else if (which == 3) out[i] = acos_poly(x[i]);
    else bz_f(x[i], out[i], out2[i]);
  }
}

}  // extern "C"

extern "C" {

// One addition-theorem step of kg_step.hpp for a single third-kind
// integral: in = {sn, cn, dn, (E-F)/m, Pi} at u and at v; out = the same at
// u + v.  Characteristics the walk transforms (n > 1) are converted on the
// way in and out as the sample loops do.  Returns 1 when the step is taken,
// 0 when it is refused.
int kgh_walk_step(const double* at_u, const double* at_v, double u, double v, double m, double n, double* out) {
  WalkState<double, 1> a, r;
  WalkStep<double, 1> w;
  const double nn[1] = {n};
  walk_init(w, m, nn);
  a.s = at_u[0]; a.c = at_u[1]; a.d = at_u[2]; a.D = at_u[3];
  walk_enter(w, 0, at_u[4], u, a.s, a.c, a.d, a);
  w.b.s = at_v[0]; w.b.c = at_v[1]; w.b.d = at_v[2]; w.b.D = at_v[3];
  walk_enter(w, 0, at_v[4], v, w.b.s, w.b.c, w.b.d, w.b);
  const bool ok = walk_step<double, 1, true>(a, w, r);
  out[0] = r.s; out[1] = r.c; out[2] = r.d; out[3] = r.D;
  out[4] = walk_pi(w, 0, r, u + v);
  return ok ? 1 : 0;
}

}  // extern "C"
