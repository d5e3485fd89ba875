// HIP kernels (gfx950) and C-ABI for the equatorial-emission image and its
// tangent (vector-Jacobian product) with respect to (a, th_o, p).
//
//   kg_eq_image    one lane per pixel: r_equatorial for subrings
//                  0..mbar_max, Keplerian redshift, g^3 r_s^-p, image sum
//   kg_eq_tangent  one lane per pixel: the same computation on
//                  kgd::Dual<3> (forward-mode, three tangents), so each lane
//                  recomputes its geodesic instead of storing samples; the
//                  lane's dI/d(a, th_o, p) is weighted by the incoming
//                  gradient and reduced in LDS to one partial per block
//   kg_eq_reduce   one block: sums the per-block partials in a fixed order
//                  (deterministic; no floating-point atomics)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/kgeo.h"
#include "kg_dual.hpp"
#include "kg_equatorial.hpp"
#include "kg_lensing.hpp"

using namespace kg;
using Dual3 = kgd::Dual<3>;

namespace {

thread_local char g_err_eq[512] = "";

constexpr int kBlock = 256;
// one wave per workgroup for the kernels whose lanes differ in cost (root
// finds of varying length, subrings that miss the disk): a CU frees a
// workgroup's slots only when its slowest wave ends
constexpr int kLaneBlock = 64;

struct EqParams {
  RayScalars<double> c;
  double p_emis;
  double ri;
  int64_t npix;
  int mbar_max;
};

struct EqOut {
  double* image;
  double* I;
  double* g;
  double* r_s;
  double* Ir;
  double* Imax;
  int32_t* Nmax;
  uint8_t* zero_mask;
};

__global__ __launch_bounds__(kLaneBlock) void kg_eq_image(EqParams p, const double* __restrict__ alpha,
                                                      const double* __restrict__ beta, EqOut o) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kLaneBlock + threadIdx.x;
  if (i >= p.npix) return;
  EqPixel<double> e;
  equatorial_pixel(p.c, alpha[i], beta[i], p.p_emis, p.ri, p.mbar_max, e);
  if (o.image) o.image[i] = e.image;
  for (int m = 0; m <= p.mbar_max; ++m) {
    const int64_t q = static_cast<int64_t>(m) * p.npix + i;
    if (o.I) o.I[q] = e.I[m];
    if (o.g) o.g[q] = e.g[m];
    if (o.r_s) o.r_s[q] = e.r_s[m];
    if (o.Ir) o.Ir[q] = e.Ir[m];
  }
  if (o.Imax) o.Imax[i] = e.Imax;
  if (o.Nmax) o.Nmax[i] = e.Nmax;
  if (o.zero_mask) o.zero_mask[i] = static_cast<uint8_t>(e.zero_mask);
}

struct TanParams {
  double a, th_o, r_o, p_emis;
  int64_t npix;
  int mbar_max;
};

__global__ __launch_bounds__(kBlock) void kg_eq_tangent(TanParams p, const double* __restrict__ alpha,
                                                        const double* __restrict__ beta,
                                                        const double* __restrict__ w, double* __restrict__ jac,
                                                        double* __restrict__ partials) {
  __shared__ double red[3][kBlock];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;
  if (i < p.npix) {
    const Dual3 a = Dual3::seed(p.a, 0);
    const Dual3 th = Dual3::seed(p.th_o, 1);
    const Dual3 pe = Dual3::seed(p.p_emis, 2);
    RayScalars<Dual3> c = make_scalars(a, th, p.r_o, 1.0);
    const Dual3 ri = r_isco(a);
    EqPixel<Dual3> e;
    equatorial_pixel(c, Dual3(alpha[i]), Dual3(beta[i]), pe, ri, p.mbar_max, e);
    if (jac) {
      jac[i] = e.image.d[0];
      jac[p.npix + i] = e.image.d[1];
      jac[2 * p.npix + i] = e.image.d[2];
    }
    const double wi = w ? w[i] : 1.0;
    // zero weight -> zero contribution, also where the tangent is not finite
    if (wi != 0.0) {
      c0 = wi * e.image.d[0];
      c1 = wi * e.image.d[1];
      c2 = wi * e.image.d[2];
    }
  }
  red[0][threadIdx.x] = c0;
  red[1][threadIdx.x] = c1;
  red[2][threadIdx.x] = c2;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] += red[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[3 * static_cast<int64_t>(blockIdx.x) + 0] = red[0][0];
    partials[3 * static_cast<int64_t>(blockIdx.x) + 1] = red[1][0];
    partials[3 * static_cast<int64_t>(blockIdx.x) + 2] = red[2][0];
  }
}

__global__ __launch_bounds__(kBlock) void kg_eq_reduce(const double* __restrict__ partials, int64_t nparts,
                                                       double* __restrict__ grad3) {
  __shared__ double red[3][kBlock];
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;
  for (int64_t b = threadIdx.x; b < nparts; b += kBlock) {
    c0 += partials[3 * b + 0];
    c1 += partials[3 * b + 1];
    c2 += partials[3 * b + 2];
  }
  red[0][threadIdx.x] = c0;
  red[1][threadIdx.x] = c1;
  red[2][threadIdx.x] = c2;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] += red[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    grad3[0] = red[0][0];
    grad3[1] = red[1][0];
    grad3[2] = red[2][0];
  }
}

// One lane per (curve k, angle i): rho_of_req_single (equatorial_lensing.py:249-262)
// for curve k's spin, inclination, target radius and subring at varphi[i].
__global__ __launch_bounds__(kLaneBlock) void kg_rho_of_req(const kgeo_ray_params* __restrict__ curves,
                                                        const double* __restrict__ req,
                                                        const int32_t* __restrict__ mbar, int64_t ncurves,
                                                        const double* __restrict__ varphi, int64_t nphi,
                                                        double* __restrict__ rho, int32_t* __restrict__ status) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kLaneBlock + threadIdx.x;
  if (g >= ncurves * nphi) return;
  const int64_t k = g / nphi, i = g - (g / nphi) * nphi;
  const kgeo_ray_params& p = curves[k];
  RayScalars<double> c = make_scalars(p.a, p.th_o, p.r_o, p.maxtaufrac);
  c.a2 = p.a2;
  c.sin_th = p.sin_th;
  c.cos_th = p.cos_th;
  c.cos2_th = p.cos2_th;
  c.rplus = p.rplus;
  c.rminus = p.rminus;
  double r;
  const int st = rho_of_req_at(c, req[k], mbar[k], varphi[i], &r);
  rho[g] = r;
  if (status) status[g] = st;
}

// One lane per photon-orbit radius: critical_curve (equatorial_lensing.py:264-293).
__global__ __launch_bounds__(kBlock) void kg_critical_curve(double a, double sin_th, double tan_th, double cos_th,
                                                            const double* __restrict__ r, int64_t n,
                                                            double* __restrict__ alpha, double* __restrict__ beta) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  critical_point(a, sin_th, tan_th, cos_th, r[i], &alpha[i], &beta[i]);
}

int eq_fail(int code, const char* msg) {
  std::snprintf(g_err_eq, sizeof(g_err_eq), "%s", msg);
  return code;
}

int eq_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::snprintf(g_err_eq, sizeof(g_err_eq), "%s launch failed: %d", what, static_cast<int>(e));
    return KGEO_E_HIP;
  }
  return KGEO_OK;
}

int vjp_check(const kgeo_ray_params* p, const double* alpha, const double* beta, int mbar_max, double* grad3) {
  if (!p) return eq_fail(KGEO_E_ARG, "params is NULL");
  if (p->npix < 0) return eq_fail(KGEO_E_ARG, "npix < 0");
  if (mbar_max < 0 || mbar_max >= kMaxSubrings) return eq_fail(KGEO_E_ARG, "mbar_max outside [0, 3]");
  if (!grad3) return eq_fail(KGEO_E_ARG, "grad3 is NULL");
  if (p->npix > 0 && (!alpha || !beta)) return eq_fail(KGEO_E_ARG, "alpha/beta is NULL");
  return KGEO_OK;
}

// The tangent kernel writes one partial triple per block into `partials`
// (every block writes its own, so nothing needs clearing first); the reduce
// kernel sums them in a fixed order into grad3.
int run_vjp(const kgeo_ray_params* p, const double* alpha, const double* beta, int mbar_max, double emis_p,
            const double* grad_image, double* jac, double* grad3, double* partials, hipStream_t st) {
  TanParams tp{p->a, p->th_o, p->r_o, emis_p, p->npix, mbar_max};
  const int64_t nb = (p->npix + kBlock - 1) / kBlock;
  if (nb > 0) {
    hipLaunchKernelGGL(kg_eq_tangent, dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, st, tp, alpha, beta,
                       grad_image, jac, partials);
    int rc = eq_launch("kg_eq_tangent");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(kg_eq_reduce, dim3(1), dim3(kBlock), 0, st, partials, nb, grad3);
  return eq_launch("kg_eq_reduce");
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of kgeo_equatorial_image_vjp_noalloc: the block partials live in
// this file-scope device array instead of a caller workspace.  Each GPU the
// library runs on holds its own copy (768 KiB) until the process ends.  Host
// code never uses the array's name as an address: scratch_ptr() asks the
// runtime for the device address once per device (with that device current)
// and keeps it.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kScratchBlocks = 32768;  // partial triples: 32768 blocks of 256 pixels
static_assert(kScratchBlocks * kBlock == KGEO_VJP_NOALLOC_MAX_NPIX, "scratch size and the header's limit differ");
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[3 * kScratchBlocks];

namespace {
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_ptr(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return eq_fail(KGEO_E_HIP, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return eq_fail(KGEO_E_ARG, "device ordinal outside [0, 63]");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return eq_fail(KGEO_E_HIP, "hipGetSymbolAddress(scratch) failed");
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch[dev];
  return KGEO_OK;
}
}  // namespace

extern "C" {

const char* kgeo_equatorial_last_error(void) { return g_err_eq; }

size_t kgeo_equatorial_vjp_workspace_bytes(int64_t npix) {
  if (npix < 0) return 0;
  const int64_t nb = (npix + kBlock - 1) / kBlock;
  return static_cast<size_t>(nb) * 3 * sizeof(double) + 256;
}

int kgeo_equatorial_image(const kgeo_ray_params* p, const double* alpha, const double* beta, int mbar_max,
                          double emis_p, double* image, double* I_sub, double* g_sub, double* r_s_sub,
                          double* Ir_sub, double* Imax, int32_t* Nmax, uint8_t* zero_mask, void* stream) {
  if (!p) return eq_fail(KGEO_E_ARG, "params is NULL");
  if (p->npix < 0) return eq_fail(KGEO_E_ARG, "npix < 0");
  if (mbar_max < 0 || mbar_max >= kMaxSubrings) return eq_fail(KGEO_E_ARG, "mbar_max outside [0, 3]");
  if (p->npix == 0) return KGEO_OK;
  if (!alpha || !beta) return eq_fail(KGEO_E_ARG, "alpha/beta is NULL");
  EqParams ep;
  ep.c = make_scalars(p->a, p->th_o, p->r_o, p->maxtaufrac);
  ep.c.a2 = p->a2;
  ep.c.sin_th = p->sin_th;
  ep.c.cos_th = p->cos_th;
  ep.c.cos2_th = p->cos2_th;
  ep.c.rplus = p->rplus;
  ep.c.rminus = p->rminus;
  ep.p_emis = emis_p;
  ep.ri = r_isco(p->a);
  ep.npix = p->npix;
  ep.mbar_max = mbar_max;
  EqOut o{image, I_sub, g_sub, r_s_sub, Ir_sub, Imax, Nmax, zero_mask};
  const unsigned nb = static_cast<unsigned>((p->npix + kLaneBlock - 1) / kLaneBlock);
  hipLaunchKernelGGL(kg_eq_image, dim3(nb), dim3(kLaneBlock), 0, static_cast<hipStream_t>(stream), ep, alpha, beta, o);
  return eq_launch("kg_eq_image");
}

int kgeo_rho_of_req(const kgeo_ray_params* curves, const double* req, const int32_t* mbar, int64_t ncurves,
                    const double* varphi, int64_t nphi, double* rho, int32_t* status, void* stream) {
  if (ncurves < 0 || nphi < 0) return eq_fail(KGEO_E_ARG, "negative size");
  if (ncurves == 0 || nphi == 0) return KGEO_OK;
  if (!curves || !req || !mbar || !varphi || !rho) return eq_fail(KGEO_E_ARG, "NULL array");
  if (ncurves > (int64_t(1) << 40) / nphi) return eq_fail(KGEO_E_ARG, "ncurves * nphi too large");
  const int64_t n = ncurves * nphi;
  const unsigned nb = static_cast<unsigned>((n + kLaneBlock - 1) / kLaneBlock);
  hipLaunchKernelGGL(kg_rho_of_req, dim3(nb), dim3(kLaneBlock), 0, static_cast<hipStream_t>(stream), curves, req, mbar,
                     ncurves, varphi, nphi, rho, status);
  return eq_launch("kg_rho_of_req");
}

int kgeo_critical_curve(const kgeo_ray_params* p, const double* r, int64_t n, double* alpha, double* beta,
                        void* stream) {
  if (!p) return eq_fail(KGEO_E_ARG, "params is NULL");
  if (n < 0) return eq_fail(KGEO_E_ARG, "n < 0");
  if (n == 0) return KGEO_OK;
  if (!r || !alpha || !beta) return eq_fail(KGEO_E_ARG, "NULL array");
  const unsigned nb = static_cast<unsigned>((n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_critical_curve, dim3(nb), dim3(kBlock), 0, static_cast<hipStream_t>(stream), p->a,
                     p->sin_th, std::tan(p->th_o), p->cos_th, r, n, alpha, beta);
  return eq_launch("kg_critical_curve");
}

int kgeo_equatorial_image_vjp(const kgeo_ray_params* p, const double* alpha, const double* beta, int mbar_max,
                              double emis_p, const double* grad_image, double* jac, double* grad3, void* workspace,
                              size_t workspace_bytes, void* stream) {
  int rc = vjp_check(p, alpha, beta, mbar_max, grad3);
  if (rc) return rc;
  if (!workspace || workspace_bytes < kgeo_equatorial_vjp_workspace_bytes(p->npix))
    return eq_fail(KGEO_E_WORKSPACE, "workspace too small");
  return run_vjp(p, alpha, beta, mbar_max, emis_p, grad_image, jac, grad3, static_cast<double*>(workspace),
                 static_cast<hipStream_t>(stream));
}

int kgeo_equatorial_image_vjp_noalloc(const kgeo_ray_params* p, const double* alpha, const double* beta,
                                      int mbar_max, double emis_p, const double* grad_image, double* jac,
                                      double* grad3, void* stream) {
  int rc = vjp_check(p, alpha, beta, mbar_max, grad3);
  if (rc) return rc;
  if (p->npix > KGEO_VJP_NOALLOC_MAX_NPIX)
    return eq_fail(KGEO_E_ARG, "npix exceeds the allocation-free scratch (8,388,608 pixels); "
                               "use kgeo_equatorial_image_vjp with a workspace");
  double* partials = nullptr;
  rc = scratch_ptr(&partials);
  if (rc) return rc;
  return run_vjp(p, alpha, beta, mbar_max, emis_p, grad_image, jac, grad3, partials,
                 static_cast<hipStream_t>(stream));
}

}  // extern "C"
