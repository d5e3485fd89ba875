// HIP kernels (gfx950) and the C-ABI of include/kgeo.h for the analytic
// Kerr raytracer.
//
// Pipeline of one kgeo_raytrace_ana call, all on the caller's stream:
//   kg_clear_flags   one wave: zero the image-wide flag word in the workspace
//   kg_setup         one lane per pixel: lambda, eta, u_+-, roots, labels,
//                    tau_tot, nmax_eq and the per-pixel constants of the
//                    theta and radial integrals -> slot-major plan [PL_N][npix];
//                    wave-ballot OR of the image-wide flags the reference's
//                    numpy code decides globally (any eta > 0, any zero step)
//   kg_theta,        one lane per (pixel, sample chunk): loads its plan slots
//   kg_radial        once (coalesced), builds its Jacobi ladder, loops over
//                    its Mino-time samples and stores r, theta, phi, t, sigma
//                    at [j][npix] -- 64 lanes write 512 contiguous bytes.
// The work is FP64 VALU (no MFMA): Carlson duplication, Landen ladders and
// transcendentals per sample.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/kgeo.h"
#include "kg_geodesic.hpp"

using namespace kg;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, const char* a = "", long long b = 0) {
  std::snprintf(g_err, sizeof(g_err), fmt, a, b);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(KGEO_E_HIP, "%s launch failed: %lld", what, static_cast<long long>(e));
  return KGEO_OK;
}

constexpr int kBlock = 256;
constexpr size_t kFlagBytes = 256;
enum : uint32_t { GF_ANY_ETA_POS = 1u, GF_ANY_STEP_ZERO = 2u };

struct DevParams {
  RayScalars<double> c;
  int64_t npix, ngeo;
  int do_phi_t;
};

DevParams make_dev(const kgeo_ray_params* p, int64_t ngeo) {
  DevParams d;
  d.c = make_scalars(p->a, p->th_o, p->r_o, p->maxtaufrac);
  d.c.a2 = p->a2;
  d.c.sin_th = p->sin_th;
  d.c.cos_th = p->cos_th;
  d.c.cos2_th = p->cos2_th;
  d.c.rplus = p->rplus;
  d.c.rminus = p->rminus;
  d.npix = p->npix;
  d.ngeo = ngeo;
  d.do_phi_t = (p->flags & KGEO_DO_PHI_T) ? 1 : 0;
  return d;
}

__device__ inline void wave_or(uint32_t* word, bool pred, uint32_t bit) {
  const unsigned long long b = __ballot(pred ? 1 : 0);
  if (b != 0ull && __lane_id() == 0) atomicOr(word, bit);
}

__global__ void kg_clear_flags(uint32_t* flags) {
  if (threadIdx.x < kFlagBytes / sizeof(uint32_t)) flags[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(kBlock) void kg_setup(DevParams p, const double* __restrict__ alpha,
                                                   const double* __restrict__ beta, double* __restrict__ plan,
                                                   uint32_t* __restrict__ gflags, int8_t* __restrict__ rclass,
                                                   int8_t* __restrict__ thclass, int32_t* __restrict__ nmax_eq,
                                                   double* __restrict__ tau_tot, uint32_t* __restrict__ status) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  bool eta_pos = false, step_zero = false;
  if (i < p.npix) {
    double pl[PL_N];
    const double al = alpha[i], be = beta[i];
    PixelSetup<double> px = plan_pixel(p.c, al, be, p.do_phi_t != 0, pl);
#pragma unroll
    for (int s = 0; s < PL_N; ++s) plan[static_cast<int64_t>(s) * p.npix + i] = pl[s];
    eta_pos = px.eta > 0.0;
    const double stop = pl[PL_STOP];
    step_zero = (p.ngeo > 1) && (stop / static_cast<double>(p.ngeo - 1) == 0.0);
    if (rclass) rclass[i] = static_cast<int8_t>(px.rclass);
    if (thclass) thclass[i] = static_cast<int8_t>(px.thclass);
    if (tau_tot) tau_tot[i] = px.tau_tot;
    if (nmax_eq) {
      // the last grid value is `stop` for ngeo > 1 in both linspace branches
      const double tl = (p.ngeo > 1) ? stop : 0.0 * stop;
      nmax_eq[i] = nmax_equatorial_at(p.c, al, be, tl);
    }
    if (status) status[i] = px.status;
  }
  wave_or(gflags, eta_pos, GF_ANY_ETA_POS);
  wave_or(gflags, step_zero, GF_ANY_STEP_ZERO);
}

// ---------------------------------------------------------------------------
// Sample kernels.  The theta kernel writes theta and parks G_phi, G_t in the
// phi and t outputs; the radial kernel then writes r and combines
// phi = I_phi + lambda G_phi, t = I_t + a^2 G_t, sigma = I_sigma + a^2 G_t
// (kerr_raytracing_ana.py:186-188).  Each kernel runs one loop specialised at
// compile time for the pixel's motion class / radial case, so a wave only
// holds the constants of its own case in registers.
// ---------------------------------------------------------------------------
struct SampleArgs {
  const double* plan;
  const uint32_t* gflags;
  const double* tau_in;
  double* tau_out;
  double* t;
  double* r;
  double* th;
  double* ph;
  double* sig;
  uint32_t* status;
  int chunks;
};

__device__ inline double tau_at(const DevParams& p, const SampleArgs& a, int64_t j, int64_t o, double stop,
                                bool any_step_zero) {
  return a.tau_in ? a.tau_in[o] : mino_tau(j, p.ngeo, stop, any_step_zero);
}

template <int KIND, bool DO_PHI_T>
__device__ inline bool theta_loop(const DevParams& p, const SampleArgs& a, const ThetaConsts<double>& tc,
                                  const JacobiLadder<double>& lad, int64_t i, int64_t j0, int64_t j1, double stop,
                                  bool any_step_zero) {
  bool bad = false;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t o = j * p.npix + i;
    const double tau = tau_at(p, a, j, o, stop, any_step_zero);
    double th = 0.0, Gph = 0.0, Gt = 0.0;
    if (KIND != THK_NONE) theta_eval<double, KIND>(tc, lad, tau, DO_PHI_T, th, Gph, Gt);
    if (a.tau_out) a.tau_out[o] = tau;
    if (a.th) a.th[o] = th;
    bad |= !isfinite(th);
    if (DO_PHI_T) {
      a.ph[o] = Gph;
      a.t[o] = Gt;
    }
  }
  return bad;
}

template <bool DO_PHI_T>
__global__ __launch_bounds__(kBlock) void kg_theta(DevParams p, SampleArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const int64_t G = p.ngeo;
  const int64_t j0 = (G * blockIdx.y) / a.chunks;
  const int64_t j1 = (G * (blockIdx.y + 1)) / a.chunks;
  ThetaConsts<double> tc;
#pragma unroll
  for (int s = 0; s < TH_NSLOT; ++s) tc.v[s] = a.plan[static_cast<int64_t>(PL_TH + s) * p.npix + i];
  const int kind = static_cast<int>(a.plan[static_cast<int64_t>(PL_THK) * p.npix + i]);
  const bool eta0 = a.plan[static_cast<int64_t>(PL_ETA0) * p.npix + i] != 0.0;
  const double stop = a.plan[static_cast<int64_t>(PL_STOP) * p.npix + i];
  const uint32_t gf = a.gflags[0];
  const bool any_eta_pos = (gf & GF_ANY_ETA_POS) != 0;
  const bool any_step_zero = (gf & GF_ANY_STEP_ZERO) != 0;
  const int k = (eta0 && !any_eta_pos) ? THK_NONE : kind;
  JacobiLadder<double> lad;
  jacobi_ladder(tc.v[TH_mp], lad);
  bool bad;
  switch (k) {
    case THK_ORD: bad = theta_loop<THK_ORD, DO_PHI_T>(p, a, tc, lad, i, j0, j1, stop, any_step_zero); break;
    case THK_VORT: bad = theta_loop<THK_VORT, DO_PHI_T>(p, a, tc, lad, i, j0, j1, stop, any_step_zero); break;
    case THK_ORD_K0: bad = theta_loop<THK_ORD_K0, DO_PHI_T>(p, a, tc, lad, i, j0, j1, stop, any_step_zero); break;
    default: bad = theta_loop<THK_NONE, DO_PHI_T>(p, a, tc, lad, i, j0, j1, stop, any_step_zero); break;
  }
  if (a.status && (bad || lad.capped)) atomicOr(&a.status[i], (bad ? ST_NONFINITE : 0u) | (lad.capped ? ST_LADDER_CAP : 0u));
}

template <int CASE, bool DO_PHI_T>
__device__ inline bool radial_loop(const DevParams& p, const SampleArgs& a, const RadialConsts<double>& rc,
                                   const JacobiLadder<double>& lad, const CombineConsts<double>& cmb, int64_t i,
                                   int64_t j0, int64_t j1, double stop, bool any_step_zero) {
  bool bad = false;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t o = j * p.npix + i;
    const double tau = tau_at(p, a, j, o, stop, any_step_zero);
    RadialOut<double> ro = radial_eval<double, CASE>(rc, lad, tau, DO_PHI_T);
    if (a.r) a.r[o] = ro.r;
    bad |= !isfinite(ro.r);
    if (DO_PHI_T) {
      const double Gph = a.ph[o];
      const double Gt = a.t[o];
      Sample<double> s;
      combine(cmb, tau, ro, 0.0, Gph, Gt, s);
      a.t[o] = s.t;
      a.ph[o] = s.ph;
      if (a.sig) a.sig[o] = s.sig;
      bad |= !(isfinite(s.t) && isfinite(s.ph) && isfinite(s.sig));
    }
  }
  return bad;
}

template <bool DO_PHI_T>
__global__ __launch_bounds__(kBlock) void kg_radial(DevParams p, SampleArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const int64_t G = p.ngeo;
  const int64_t j0 = (G * blockIdx.y) / a.chunks;
  const int64_t j1 = (G * (blockIdx.y + 1)) / a.chunks;
  RadialConsts<double> rc;
#pragma unroll
  for (int s = 0; s < R_NSLOT; ++s) rc.v[s] = a.plan[static_cast<int64_t>(PL_R + s) * p.npix + i];
  rc.rcase = static_cast<int>(a.plan[static_cast<int64_t>(PL_RCASE) * p.npix + i]);
  const double lam = a.plan[static_cast<int64_t>(PL_LAM) * p.npix + i];
  const double stop = a.plan[static_cast<int64_t>(PL_STOP) * p.npix + i];
  const bool any_step_zero = (a.gflags[0] & GF_ANY_STEP_ZERO) != 0;
  const CombineConsts<double> cmb = combine_consts(p.c, lam);
  JacobiLadder<double> lad;
  jacobi_ladder(rc.v[R_k], lad);
  bool bad;
  switch (rc.rcase) {
    case 4: bad = radial_loop<4, DO_PHI_T>(p, a, rc, lad, cmb, i, j0, j1, stop, any_step_zero); break;
    case 3: bad = radial_loop<3, DO_PHI_T>(p, a, rc, lad, cmb, i, j0, j1, stop, any_step_zero); break;
    default: bad = radial_loop<2, DO_PHI_T>(p, a, rc, lad, cmb, i, j0, j1, stop, any_step_zero); break;
  }
  if (a.status && (bad || lad.capped)) atomicOr(&a.status[i], (bad ? ST_NONFINITE : 0u) | (lad.capped ? ST_LADDER_CAP : 0u));
}

__global__ __launch_bounds__(kBlock) void kg_ellip_pi(const double* __restrict__ n, const double* __restrict__ phi,
                                                      const double* __restrict__ m, double* __restrict__ out,
                                                      int64_t count) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= count) return;
  const double nn = n[i], mm = m[i];
  Amp<double> A = amp_from_angle(phi[i]);
  LegendreBase<double> L = legendre_base(A, mm);
  double Pc = 0.0;
  if (A.j != 0.0) Pc = ellip_pi_complete(nn, mm, ellip_K(mm));
  // phi itself for |phi| <= 1e-10 (scipy_ellip_binding.py:126-128)
  out[i] = (kfabs(phi[i]) <= 1.0e-10) ? phi[i] : ellip_pi_amp(L, nn, mm, Pc);
}

// Per-sample integer / fractional counts (kerr_raytracing_utils.py:551-684).
__global__ __launch_bounds__(kBlock) void kg_counts(DevParams p, const double* __restrict__ alpha,
                                                    const double* __restrict__ beta, const double* __restrict__ tau,
                                                    int64_t ntau, double* __restrict__ n_eq,
                                                    double* __restrict__ n_pol) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const RayScalars<double>& c = p.c;
  const double al = alpha[i], be = beta[i];
  const double lam = -al * c.sin_th;
  const double eta = (al * al - c.a2) * c.cos2_th + be * be;
  const bool vort = eta <= 0.0;
  double up = 0.0, um = 0.0, ratio = 0.0, a2um = 0.0, F0 = 0.0, Kinc = 0.0, K = 0.0, w = 0.0;
  if (!vort) {
    uplus_uminus(c, lam, eta, up, um);
    if (c.minspin) {
      ratio = 0.0 * um;
      a2um = -(eta + lam * lam);
    } else {
      ratio = up / um;
      a2um = c.a2 * um;
    }
    K = ellip_K(ratio);
    Kinc = K;
    F0 = legendre_F(legendre_base(amp_from_sin(c.cos_th / ksqrt(up)), ratio), K);
    w = ksqrt(-a2um);
  }
  for (int64_t j = 0; j < ntau; ++j) {
    const int64_t o = j * p.npix + i;
    const double tj = tau[o];
    if (vort) {
      if (n_eq) n_eq[o] = -2.0;
      if (n_pol) n_pol[o] = -2.0;
      continue;
    }
    const double x = tj * w;
    if (n_eq) {
      double n;
      if (c.h > 0.0) n = (be <= 0.0) ? kfloor((x - F0) / (2.0 * Kinc)) : kfloor((x + F0) / (2.0 * Kinc)) - 1.0;
      else n = (be >= 0.0) ? kfloor((x + F0) / (2.0 * Kinc)) : kfloor((x - F0) / (2.0 * Kinc)) - 1.0;
      n_eq[o] = n;
    }
    if (n_pol) n_pol[o] = (w * tj) / (4.0 * K);
  }
}

// Samples per lane chunking: enough (pixel, chunk) lanes to cover the chip
// (256 CUs x 8 waves of 64) while keeping chunks >= 16 samples.
int pick_chunks(int64_t npix, int64_t ngeo) {
  const int64_t target_lanes = 256LL * 8 * 64;
  int64_t c = (target_lanes + npix - 1) / (npix > 0 ? npix : 1);
  if (c < 1) c = 1;
  const int64_t max_c = ngeo / 16 > 0 ? ngeo / 16 : 1;
  if (c > max_c) c = max_c;
  if (c > 65535) c = 65535;
  return static_cast<int>(c);
}

int env_chunks() {
  const char* s = std::getenv("KGEO_CHUNKS");
  return s ? std::atoi(s) : 0;
}

int validate(const kgeo_ray_params* p, const double* alpha, const double* beta, void* ws, size_t wsb) {
  if (!p) return fail(KGEO_E_ARG, "params is NULL%s%lld");
  if (p->npix < 0) return fail(KGEO_E_ARG, "npix < 0%s (%lld)", "", p->npix);
  if (p->npix > 0 && (!alpha || !beta)) return fail(KGEO_E_ARG, "alpha/beta is NULL%s%lld");
  if (p->npix > 0 && !ws) return fail(KGEO_E_WORKSPACE, "workspace is NULL%s%lld");
  if (wsb < kgeo_workspace_bytes(p->npix))
    return fail(KGEO_E_WORKSPACE, "workspace too small%s: need %lld bytes", "",
                static_cast<long long>(kgeo_workspace_bytes(p->npix)));
  if (reinterpret_cast<uintptr_t>(ws) % 256 != 0) return fail(KGEO_E_WORKSPACE, "workspace not 256-byte aligned%s%lld");
  return KGEO_OK;
}

int run_plan(const kgeo_ray_params* p, const double* alpha, const double* beta, int8_t* rclass, int8_t* thclass,
             int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* ws, hipStream_t st) {
  DevParams d = make_dev(p, p->ngeo);
  if (d.npix == 0) return KGEO_OK;
  uint32_t* gflags = static_cast<uint32_t*>(ws);
  double* plan = reinterpret_cast<double*>(static_cast<char*>(ws) + kFlagBytes);
  hipLaunchKernelGGL(kg_clear_flags, dim3(1), dim3(64), 0, st, gflags);
  int rc = check_launch("kg_clear_flags");
  if (rc) return rc;
  const unsigned nb = static_cast<unsigned>((d.npix + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_setup, dim3(nb), dim3(kBlock), 0, st, d, alpha, beta, plan, gflags, rclass, thclass, nmax_eq,
                     tau_tot, status);
  return check_launch("kg_setup");
}

int run_sample(const kgeo_ray_params* p, int64_t ngeo, const double* tau_in, double* tau, double* t, double* r,
               double* th, double* ph, double* sig, uint32_t* status, const void* ws, hipStream_t st) {
  DevParams d = make_dev(p, ngeo);
  if (d.npix == 0 || ngeo <= 0) return KGEO_OK;
  if (d.do_phi_t && (!t || !ph))
    return fail(KGEO_E_ARG, "t and ph outputs are required with KGEO_DO_PHI_T%s%lld");
  const uint32_t* gflags = static_cast<const uint32_t*>(ws);
  const double* plan = reinterpret_cast<const double*>(static_cast<const char*>(ws) + kFlagBytes);
  const unsigned nb = static_cast<unsigned>((d.npix + kBlock - 1) / kBlock);
  int chunks = env_chunks();
  if (chunks <= 0) chunks = pick_chunks(d.npix, ngeo);
  if (chunks > ngeo) chunks = static_cast<int>(ngeo);
  dim3 grid(nb, static_cast<unsigned>(chunks));
  SampleArgs a{plan, gflags, tau_in, tau, t, r, th, ph, sig, status, chunks};
  if (d.do_phi_t) {
    hipLaunchKernelGGL(kg_theta<true>, grid, dim3(kBlock), 0, st, d, a);
    int rc = check_launch("kg_theta");
    if (rc) return rc;
    hipLaunchKernelGGL(kg_radial<true>, grid, dim3(kBlock), 0, st, d, a);
  } else {
    hipLaunchKernelGGL(kg_theta<false>, grid, dim3(kBlock), 0, st, d, a);
    int rc = check_launch("kg_theta");
    if (rc) return rc;
    hipLaunchKernelGGL(kg_radial<false>, grid, dim3(kBlock), 0, st, d, a);
  }
  return check_launch("kg_radial");
}

}  // namespace

extern "C" {

int kgeo_abi_version(void) { return KGEO_ABI_VERSION; }

const char* kgeo_last_error(void) { return g_err; }

void kgeo_fill_params(kgeo_ray_params* p) {
  if (!p) return;
  p->a2 = p->a * p->a;
  p->sin_th = std::sin(p->th_o);
  p->cos_th = std::cos(p->th_o);
  p->cos2_th = p->cos_th * p->cos_th;
  const double root = std::sqrt(1.0 - p->a2);
  p->rplus = 1.0 + root;
  p->rminus = 1.0 - root;
}

size_t kgeo_workspace_bytes(int64_t npix) {
  if (npix < 0) return 0;
  return kFlagBytes + static_cast<size_t>(npix) * PL_N * sizeof(double);
}

int kgeo_plan(const kgeo_ray_params* p, const double* alpha, const double* beta, int8_t* rclass, int8_t* thclass,
              int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace, size_t workspace_bytes,
              void* stream) {
  int rc = validate(p, alpha, beta, workspace, workspace_bytes);
  if (rc) return rc;
  return run_plan(p, alpha, beta, rclass, thclass, nmax_eq, tau_tot, status, workspace,
                  static_cast<hipStream_t>(stream));
}

int kgeo_sample(const kgeo_ray_params* p, const double* tau_in, int64_t ntau, double* tau, double* t, double* r,
                double* th, double* ph, double* sig, uint32_t* status, const void* workspace, size_t workspace_bytes,
                void* stream) {
  if (!p) return fail(KGEO_E_ARG, "params is NULL%s%lld");
  if (p->npix < 0 || ntau < 0) return fail(KGEO_E_ARG, "negative size%s%lld");
  if (p->npix > 0 && !workspace) return fail(KGEO_E_WORKSPACE, "workspace is NULL%s%lld");
  if (workspace_bytes < kgeo_workspace_bytes(p->npix)) return fail(KGEO_E_WORKSPACE, "workspace too small%s%lld");
  return run_sample(p, ntau, tau_in, tau, t, r, th, ph, sig, status, workspace, static_cast<hipStream_t>(stream));
}

int kgeo_raytrace_ana(const kgeo_ray_params* p, const double* alpha, const double* beta, double* tau, double* t,
                      double* r, double* th, double* ph, double* sig, int8_t* rclass, int8_t* thclass,
                      int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace, size_t workspace_bytes,
                      void* stream) {
  int rc = validate(p, alpha, beta, workspace, workspace_bytes);
  if (rc) return rc;
  if (p->ngeo < 1) return fail(KGEO_E_ARG, "ngeo < 1%s (%lld)", "", p->ngeo);
  hipStream_t st = static_cast<hipStream_t>(stream);
  rc = run_plan(p, alpha, beta, rclass, thclass, nmax_eq, tau_tot, status, workspace, st);
  if (rc) return rc;
  return run_sample(p, p->ngeo, nullptr, tau, t, r, th, ph, sig, status, workspace, st);
}

int kgeo_coords_at_tau(const kgeo_ray_params* p, const double* alpha, const double* beta, const double* tau_in,
                       int64_t ntau, double* t, double* r, double* th, double* ph, double* sig, uint32_t* status,
                       void* workspace, size_t workspace_bytes, void* stream) {
  int rc = validate(p, alpha, beta, workspace, workspace_bytes);
  if (rc) return rc;
  if (ntau < 0) return fail(KGEO_E_ARG, "ntau < 0%s (%lld)", "", ntau);
  if (ntau > 0 && p->npix > 0 && !tau_in) return fail(KGEO_E_ARG, "tau_in is NULL%s%lld");
  hipStream_t st = static_cast<hipStream_t>(stream);
  rc = run_plan(p, alpha, beta, nullptr, nullptr, nullptr, nullptr, status, workspace, st);
  if (rc) return rc;
  return run_sample(p, ntau, tau_in, nullptr, t, r, th, ph, sig, status, workspace, st);
}

int kgeo_crossings(const kgeo_ray_params* p, const double* alpha, const double* beta, const double* tau,
                   int64_t ntau, double* n_equatorial, double* n_poloidal, void* stream) {
  if (!p) return fail(KGEO_E_ARG, "params is NULL%s%lld");
  if (p->npix < 0 || ntau < 0) return fail(KGEO_E_ARG, "negative size%s%lld");
  if (p->npix == 0 || ntau == 0) return KGEO_OK;
  if (!alpha || !beta || !tau) return fail(KGEO_E_ARG, "NULL array%s%lld");
  DevParams d = make_dev(p, ntau);
  const unsigned nb = static_cast<unsigned>((d.npix + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_counts, dim3(nb), dim3(kBlock), 0, static_cast<hipStream_t>(stream), d, alpha, beta, tau,
                     ntau, n_equatorial, n_poloidal);
  return check_launch("kg_counts");
}

int kgeo_ellip_pi(const double* n, const double* phi, const double* m, double* out, int64_t count, void* stream) {
  if (count < 0) return fail(KGEO_E_ARG, "count < 0%s (%lld)", "", count);
  if (count == 0) return KGEO_OK;
  if (!n || !phi || !m || !out) return fail(KGEO_E_ARG, "NULL array%s%lld");
  const unsigned nb = static_cast<unsigned>((count + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_ellip_pi, dim3(nb), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, phi, m, out, count);
  return check_launch("kg_ellip_pi");
}

}  // extern "C"
