// HIP kernels (gfx950) and C-ABI for the polarized equatorial image and the
// emission models (kg_emission.hpp).
//
//   kg_stokes      one lane per pixel: r_equatorial for subrings
//                  0..mbar_max, then per emitting subring the fluid
//                  velocity, redshift, comoving-frame field and wave vector,
//                  the Penrose-Walker constant, EVPA, emissivity and the
//                  observed I, Q, U (make_image / Iobs)
//   kg_model_eval  one lane per point: u^mu, tetrad, B^i, Omega_F, b^2 and
//                  j_rest of the models (Velocity / Bfield / Emissivity)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "../../include/kgeo.h"
#include "kg_emission.hpp"

using namespace kg;

namespace {

thread_local char g_err_em[512] = "";

constexpr int kBlock = 256;
constexpr int kMaxSub = 1024;
// kg_stokes runs one wave per workgroup (pixels differ in cost: subrings
// that miss the disk or fall inside the horizon skip the models) and two
// waves per SIMD: unbounded it needs about 425 VGPR+AGPR, one wave; held to
// 256 with spills the 800^2 test image takes 0.30 instead of 0.40 ms.
constexpr int kStokesBlock = 64;

int em_fail(int code, const char* msg) {
  std::snprintf(g_err_em, sizeof(g_err_em), "%s", msg);
  return code;
}

int em_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::snprintf(g_err_em, sizeof(g_err_em), "%s launch failed: %s", what, hipGetErrorString(e));
    return KGEO_E_HIP;
  }
  return KGEO_OK;
}

bool in01(double x) { return x >= 0.0 && x <= 1.0; }

int check_field(const kgeo_bfield_model& f, const char* who) {
  if (f.kind < KGEO_BF_SIMPLE || f.kind > KGEO_BF_FROMFILE) {
    std::snprintf(g_err_em, sizeof(g_err_em), "%s: unknown field kind %d", who, f.kind);
    return KGEO_E_ARG;
  }
  if (f.kind == KGEO_BF_FROMFILE && (!f.table || f.table_n < 1)) return em_fail(KGEO_E_ARG, "fromfile field without a table");
  return KGEO_OK;
}

// The checks the reference's constructors and u_lab make
// (velocities.py:23-68, 284-292, 442-453, 527-528).
int check_model(const kgeo_emission_model& M) {
  const kgeo_velocity_model& v = M.velocity;
  if (v.kind < KGEO_VEL_ZAMO || v.kind > KGEO_VEL_DRIFTFRAME) return em_fail(KGEO_E_ARG, "unknown velocity kind");
  if ((v.kind == KGEO_VEL_SUBKEP || v.kind == KGEO_VEL_GENERAL) && !in01(v.fac_subkep))
    return em_fail(KGEO_E_ARG, "fac_subkep should be in the range [0,1]");
  if (v.kind == KGEO_VEL_GENERAL && !in01(v.beta_phi)) return em_fail(KGEO_E_ARG, "beta_phi should be in the range [0,1]");
  if (v.kind == KGEO_VEL_GENERAL && !in01(v.beta_r)) return em_fail(KGEO_E_ARG, "beta_r should be in the range [0,1]");
  if (v.kind == KGEO_VEL_FROMFILE && (!v.table || v.table_n < 1)) return em_fail(KGEO_E_ARG, "fromfile velocity without a table");
  if (v.kind == KGEO_VEL_DRIFTFRAME) {
    if (!(v.nu_parallel >= -1.0 && v.nu_parallel <= 1.0)) return em_fail(KGEO_E_ARG, "nu_parallel should be in the range (-1,1)");
    int rc = check_field(v.drift_field, "drift field");
    if (rc) return rc;
    if (!bfield_has_omega(v.drift_field.kind))
      return em_fail(KGEO_E_ARG, "driftframe needs a field with a field-line angular velocity (omega_field)");
  }
  int rc = check_field(M.bfield, "bfield");
  if (rc) return rc;
  const int ek = M.emissivity.kind;
  if (ek < KGEO_EMIS_CONSTANT || ek > KGEO_EMIS_THERMAL) return em_fail(KGEO_E_ARG, "unknown emissivity kind");
  return KGEO_OK;
}

__global__ __launch_bounds__(kStokesBlock) __attribute__((amdgpu_waves_per_eu(2))) void kg_stokes(
    RayScalars<double> c, int north, kgeo_emission_model M, const double* __restrict__ alpha,
    const double* __restrict__ beta, int64_t npix, int nsub, StokesArrays o) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kStokesBlock + threadIdx.x;
  if (i >= npix) return;
  const double bsq0 = (M.emissivity.kind == KGEO_EMIS_THERMAL)
                          ? field_bsq(M.bfield, M.velocity, c.a, M.emissivity.Rb, equator())
                          : 0.0;
  stokes_pixel(c, north != 0, M, alpha[i], beta[i], nsub, bsq0, i, o);
}

struct EvalOut {
  double *u, *tet, *B, *omega, *bsq, *j;
};

__global__ __launch_bounds__(kBlock) void kg_model_eval(kgeo_emission_model M, double a, const double* __restrict__ r,
                                                        const double* __restrict__ th, int64_t n,
                                                        const double* __restrict__ g,
                                                        const double* __restrict__ sinthb,
                                                        const double* __restrict__ bmag, EvalOut o) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const Angle an = th ? angle_of(th[i]) : equator();
  model_point(M, a, r[i], an, i, n, g, sinthb, bmag, o.u, o.tet, o.B, o.omega, o.bsq, o.j);
}

}  // namespace

extern "C" {

const char* kgeo_emission_last_error(void) { return g_err_em; }

int kgeo_equatorial_stokes(const kgeo_ray_params* p, const double* alpha, const double* beta, int32_t mbar_max,
                           const kgeo_emission_model* model, double* I, double* Q, double* U, double* g,
                           double* r_s, double* sin_thb, double* Ir, double* nmax_eq, double* n_poloidal,
                           double* lp, double* Ie, void* stream) {
  if (!p || !model) return em_fail(KGEO_E_ARG, "params or model is NULL");
  if (p->npix < 0) return em_fail(KGEO_E_ARG, "npix < 0");
  if (mbar_max < 0 || mbar_max >= kMaxSub) return em_fail(KGEO_E_ARG, "mbar_max should be an integer in [0, 1023]");
  if (!(std::fabs(p->a) < 1.0)) return em_fail(KGEO_E_ARG, "|a| should be a float in range [0,1)");
  if (!(p->r_o >= 100.0)) return em_fail(KGEO_E_ARG, "r_o should be a float > 100");
  if (!(p->th_o > 0.0 && p->th_o < kPiD && p->th_o != kHalfPiD))
    return em_fail(KGEO_E_ARG, "th_o should be a float in range (0,pi/2) or (pi/2,pi)");
  int rc = check_model(*model);
  if (rc) return rc;
  if (model->pathlength && velocity_has_theta(model->velocity))
    return em_fail(KGEO_E_ARG, "calc_pathlength assumes that u2==0!");
  if (p->npix == 0) return KGEO_OK;
  if (!alpha || !beta) return em_fail(KGEO_E_ARG, "alpha/beta is NULL");
  RayScalars<double> c = make_scalars(p->a, p->th_o, p->r_o, p->maxtaufrac);
  c.a2 = p->a2;
  c.sin_th = p->sin_th;
  c.cos_th = p->cos_th;
  c.cos2_th = p->cos2_th;
  c.rplus = p->rplus;
  c.rminus = p->rminus;
  StokesArrays o{I, Q, U, g, r_s, sin_thb, Ir, nmax_eq, n_poloidal, lp, Ie};
  const unsigned nb = static_cast<unsigned>((p->npix + kStokesBlock - 1) / kStokesBlock);
  const int north = p->th_o < kHalfPiD;
  hipLaunchKernelGGL(kg_stokes, dim3(nb), dim3(kStokesBlock), 0, static_cast<hipStream_t>(stream), c, north, *model, alpha,
                     beta, p->npix, mbar_max + 1, o);
  return em_launch("kg_stokes");
}

int kgeo_model_eval(const kgeo_emission_model* model, double a, const double* r, const double* th, int64_t n,
                    const double* g, const double* sinthb, const double* bmag, double* u, double* tetrad, double* B,
                    double* omega, double* bsq, double* j, void* stream) {
  if (!model) return em_fail(KGEO_E_ARG, "model is NULL");
  if (n < 0) return em_fail(KGEO_E_ARG, "n < 0");
  if (!(std::fabs(a) < 1.0)) return em_fail(KGEO_E_ARG, "|a| should be a float in range [0,1)");
  int rc = check_model(*model);
  if (rc) return rc;
  if (omega && !bfield_has_omega(model->bfield.kind))
    return em_fail(KGEO_E_ARG, "omega_field is not implemented for this field");
  if (j && model->emissivity.kind == KGEO_EMIS_THERMAL && (!g || !sinthb || (model->emissivity.bfield_model && !bmag)))
    return em_fail(KGEO_E_ARG, "thermal jrest needs g, sinthetab (and Bmag with bfield_model)");
  if (n == 0) return KGEO_OK;
  if (!r) return em_fail(KGEO_E_ARG, "r is NULL");
  EvalOut o{u, tetrad, B, omega, bsq, j};
  const unsigned nb = static_cast<unsigned>((n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_model_eval, dim3(nb), dim3(kBlock), 0, static_cast<hipStream_t>(stream), *model, a, r, th, n,
                     g, sinthb, bmag, o);
  return em_launch("kg_model_eval");
}

}  // extern "C"
