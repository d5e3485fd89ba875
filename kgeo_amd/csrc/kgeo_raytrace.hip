// HIP kernels (gfx950) and the C-ABI of include/kgeo.h for the analytic
// Kerr raytracer.
//
// Pipeline of one kgeo_raytrace_ana call, on the caller's stream except the
// plan's radial constants and the radial heads, far and repair passes, which
// run on a per-device side stream beside the theta work and are joined
// before the radial walk:
//   kg_clear_flags   one wave: zero the flag page of the workspace
//   kg_setup         one lane per pixel: lambda, eta, u_+-, roots, labels,
//                    tau_tot, nmax_eq and the per-pixel constants of the
//                    theta and radial integrals -> slot-major plan [PL_N][npix];
//                    wave-ballot OR of the image-wide flags the reference's
//                    numpy code decides globally (any eta > 0, any zero step);
//                    with a side stream, as kg_setup_theta (all but the
//                    radial constants) and kg_setup_radial (side stream)
// then, on the linspace grid, per coordinate (theta first: its G_phi, G_t
// are parked in the phi and t outputs; the radial stores combine them):
//   kg_*_heads       one lane per (pixel, head): direct evaluations (Landen
//                    ladder, Carlson duplication) at the steps h and L h, the
//                    last sample and each block's first sample -> records
//   kg_*_far         one lane per (pixel, block): the block's other segment
//                    starts by long addition steps -> records; a radial
//                    block whose long step is refused (a pole of a
//                    third-kind integral inside it) is queued for
//   kg_radial_repair one lane per (queued block, segment): the block's
//                    remaining segment starts by direct evaluation -> records
//   kg_*_walk        one lane per (pixel, sample chunk): every other sample
//                    by the addition theorems from the records; 64 lanes
//                    write 512 contiguous bytes per output and sample
//   kg_*_fixup       one lane per (queued range, sample): direct evaluation
//                    of the samples where the walk refused a step
// (kg_sample.hpp).  Caller-given Mino times and grids of at most two samples
// take kg_theta / kg_radial, one combined loop per coordinate.  The work is
// FP64 VALU (no MFMA).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "../../include/kgeo.h"
#include "kg_counts.hpp"
#include "kg_sample.hpp"

using namespace kg;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, const char* a = "", long long b = 0) {
  std::snprintf(g_err, sizeof(g_err), fmt, a, b);
  return code;
}

int fail_msg(int code, const char* msg) {
  std::snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(KGEO_E_HIP, "%s launch failed: %lld", what, static_cast<long long>(e));
  return KGEO_OK;
}

constexpr int kBlock = 256;
// The sample passes run one wave per workgroup: their waves differ widely in
// cost (radial case, refused steps), and a CU's slot is freed per workgroup,
// so a multi-wave workgroup would hold its SIMDs until its slowest wave ends.
// (The theta walk, whose waves cost alike, keeps kBlock.)
constexpr int kPassBlock = 64;
constexpr size_t kFlagBytes = 256;

struct DevParams {
  RayScalars<double> c;
  int64_t npix, ngeo;
  int do_phi_t;
  int north;  // th_o < pi/2 (n_equatorial_crossings' branch test)
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
  d.north = (p->th_o < kHalfPi) ? 1 : 0;
  return d;
}


__global__ void kg_clear_flags(uint32_t* flags) {
  if (threadIdx.x < kFlagBytes / sizeof(uint32_t)) flags[threadIdx.x] = 0u;
}

// Word 0 of the flag page holds the image-wide flags (kg_setup); words 1 and
// 2 count the theta and radial fixup items of one sample pass, word 3 the
// radial repair items.
constexpr int kThetaCount = 1;
constexpr int kRadialCount = 2;
constexpr int kRepairCount = 3;  // the radial far pass's repair items

__global__ void kg_clear_counters(uint32_t* flags) {
  if (threadIdx.x == 0) {
    flags[kThetaCount] = 0u;
    flags[kRadialCount] = 0u;
    flags[kRepairCount] = 0u;
  }
}

// Two waves per SIMD: unbounded the set-up needs about 300 VGPR+AGPR (one
// wave); held to 256 with spills it takes 0.23 instead of 0.27 ms at 512^2.
__global__ __launch_bounds__(kPassBlock) __attribute__((amdgpu_waves_per_eu(2))) void kg_setup(
    DevParams p, const double* __restrict__ alpha, const double* __restrict__ beta, double* __restrict__ plan,
    uint32_t* __restrict__ gflags, int8_t* __restrict__ rclass, int8_t* __restrict__ thclass,
    int32_t* __restrict__ nmax_eq, double* __restrict__ tau_tot, uint32_t* __restrict__ status) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  SetupOut so{false, false};
  if (i < p.npix)
    so = setup_pixel(p.c, p.ngeo, p.do_phi_t != 0, alpha[i], beta[i], plan, p.npix, i, rclass, thclass, nmax_eq,
                     tau_tot, status, i);
  wave_or(gflags, so.eta_pos, GF_ANY_ETA_POS);
  wave_or(gflags, so.step_zero, GF_ANY_STEP_ZERO);
}

// The plan kernel in two halves: kg_setup_theta (labels, tau_tot, flags and
// every plan slot but the radial constants) on the caller's stream, and
// kg_setup_radial (the radial constants, about two thirds of kg_setup's
// time) on the side stream, where the radial passes that read them run.
__global__ __launch_bounds__(kPassBlock) __attribute__((amdgpu_waves_per_eu(2))) void kg_setup_theta(
    DevParams p, const double* __restrict__ alpha, const double* __restrict__ beta, double* __restrict__ plan,
    uint32_t* __restrict__ gflags, int8_t* __restrict__ rclass, int8_t* __restrict__ thclass,
    int32_t* __restrict__ nmax_eq, double* __restrict__ tau_tot, uint32_t* __restrict__ status) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  SetupOut so{false, false};
  if (i < p.npix)
    so = setup_pixel<true>(p.c, p.ngeo, p.do_phi_t != 0, alpha[i], beta[i], plan, p.npix, i, rclass, thclass,
                           nmax_eq, tau_tot, status, i);
  wave_or(gflags, so.eta_pos, GF_ANY_ETA_POS);
  wave_or(gflags, so.step_zero, GF_ANY_STEP_ZERO);
}

__global__ __launch_bounds__(kPassBlock) __attribute__((amdgpu_waves_per_eu(2))) void kg_setup_radial(
    DevParams p, const double* __restrict__ alpha, const double* __restrict__ beta, double* __restrict__ plan) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  if (i < p.npix) setup_radial_pixel(p.c, p.do_phi_t != 0, alpha[i], beta[i], plan, p.npix, i);
}

// ---------------------------------------------------------------------------
// Direct sample kernels (caller-given Mino times, ngeo <= 2): grid (pixel
// blocks, sample chunks).  kg_theta writes theta and parks G_phi, G_t in ph,
// t; kg_radial writes r and combines phi, t, sigma (kg_sample.hpp).
// ---------------------------------------------------------------------------
struct SampleArgs {
  const double* plan;
  const uint32_t* gflags;
  SampleOut out;
  uint32_t* status;
  int chunks;
};

template <bool DO_PHI_T>
__global__ __launch_bounds__(kBlock) void kg_theta(DevParams p, SampleArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const int64_t j0 = chunk_start(p.ngeo, blockIdx.y, a.chunks);
  const int64_t j1 = chunk_start(p.ngeo, blockIdx.y + 1, a.chunks);
  const uint32_t st = theta_pixel<DO_PHI_T, double, false>(a.plan, p.npix, i, a.gflags[0], a.out, i, j0, j1);
  if (a.status && st) atomicOr(&a.status[i], st);
}

template <bool DO_PHI_T>
__global__ __launch_bounds__(kBlock) void kg_radial(DevParams p, SampleArgs a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const int64_t j0 = chunk_start(p.ngeo, blockIdx.y, a.chunks);
  const int64_t j1 = chunk_start(p.ngeo, blockIdx.y + 1, a.chunks);
  const uint32_t st = radial_pixel<DO_PHI_T, double, false>(p.c, a.plan, p.npix, i, a.gflags[0], a.out, i, j0, j1);
  if (a.status && st) atomicOr(&a.status[i], st);
}

// ---------------------------------------------------------------------------
// The walk in three passes per coordinate (kg_sample.hpp): anchor -> walk ->
// fixup for theta, then the same for r (whose stores combine phi, t, sigma
// with the theta parts).  The fixup queue holds one 64-bit item per queued
// range: walked << 63 | first sample << 32 | pixel.
// ---------------------------------------------------------------------------
struct FixupQueue {
  unsigned long long* items;
  uint32_t* count;
  unsigned long long cap;
};

struct DevPush {
  FixupQueue q;
  __device__ void operator()(int64_t i, int64_t j, bool walked) const {
    const uint32_t s = atomicAdd(q.count, 1u);
    if (s < q.cap)
      q.items[s] = (walked ? (1ull << 63) : 0ull) | (static_cast<unsigned long long>(j) << 32) |
                   static_cast<unsigned long long>(static_cast<uint32_t>(i));
  }
};

// The radial far pass hands a block with a lost long step to the repair
// pass through the same queue storage (the theta fixups are done with it, and
// the radial walk's pushes come after the repair pass).  At most one item per
// block and pixel: within the queue's npix * ceil(ngeo / 8) items.
struct DevRepair {
  static constexpr bool kOn = true;
  FixupQueue q;
  __device__ void operator()(int64_t i, int64_t j, bool walked) const { DevPush{q}(i, j, walked); }
};

__device__ inline void unpack_item(unsigned long long e, int64_t& i, int64_t& j, bool& walked) {
  i = static_cast<int64_t>(e & 0xffffffffull);
  j = static_cast<int64_t>((e >> 32) & 0x7fffffffull);
  walked = (e >> 63) != 0ull;
}

// grid (pixel blocks, heads): the direct evaluations of the anchor pass
template <bool DO_PHI_T>
// last = false: every head but the last sample's (head 2), which the theta
// far and walk passes do not read; last = true: head 2 alone (grid y = 1),
// which the side stream takes (it emits the last sample for the radial walk)
__global__ __launch_bounds__(kPassBlock) void kg_theta_heads(DevParams p, SampleArgs a, Records rt, int64_t nhead,
                                                             int last) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  if (i >= p.npix) return;
  uint32_t st = 0u;
  if (last) {
    st = theta_head_pixel<DO_PHI_T>(a.plan, p.npix, i, a.gflags[0], a.out, i, 2, rt);
  } else {
    for (int64_t y = blockIdx.y; y < nhead - 1; y += gridDim.y)
      st |= theta_head_pixel<DO_PHI_T>(a.plan, p.npix, i, a.gflags[0], a.out, i, y < 2 ? y : y + 1, rt);
  }
  if (a.status && st) atomicOr(&a.status[i], st);
}

// grid (pixel blocks, anchor blocks): the long steps of the anchor pass
template <bool DO_PHI_T>
__global__ __launch_bounds__(kPassBlock) void kg_theta_far(DevParams p, SampleArgs a, Records rt, int64_t nblk) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  if (i >= p.npix) return;
  uint32_t st = 0u;
  for (int64_t b = blockIdx.y; b < nblk; b += gridDim.y)
    st |= theta_far_pixel<DO_PHI_T>(a.plan, p.npix, i, a.gflags[0], a.out, i, b, rt);
  if (a.status && st) atomicOr(&a.status[i], st);
}

template <bool DO_PHI_T>
// Four waves per SIMD (128 VGPR, a few spills): 1.31 against 1.40 ms at
// three (512^2 x 500).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void kg_theta_walk(DevParams p,
    SampleArgs a, Records rt, FixupQueue q) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const int64_t j0 = chunk_start(p.ngeo, blockIdx.y, a.chunks);
  const int64_t j1 = chunk_start(p.ngeo, blockIdx.y + 1, a.chunks);
  DevPush push{q};
  const uint32_t st = theta_walk_pixel<DO_PHI_T>(a.plan, p.npix, i, a.gflags[0], a.out, i, j0, j1, rt, push);
  if (a.status && st) atomicOr(&a.status[i], st);
}

template <bool DO_PHI_T>
// one lane per (queued range, sample of the range) (grid-stride)
__global__ __launch_bounds__(kPassBlock) void kg_theta_fixup(DevParams p, SampleArgs a, FixupQueue q) {
  const unsigned long long c = *q.count;
  const unsigned long long n = (c < q.cap ? c : q.cap) * kFixupSpan;
  for (unsigned long long e = static_cast<unsigned long long>(blockIdx.x) * kPassBlock + threadIdx.x; e < n;
       e += static_cast<unsigned long long>(gridDim.x) * kPassBlock) {
    int64_t i, j;
    bool walked;
    unpack_item(q.items[e / kFixupSpan], i, j, walked);
    const int64_t sg = static_cast<int64_t>(e % kFixupSpan);
    const uint32_t st = theta_fixup_sample<DO_PHI_T>(a.plan, p.npix, i, a.gflags[0], a.out, i, j, walked, sg);
    if (a.status && st) atomicOr(&a.status[i], st);
  }
}

template <bool DO_PHI_T>
__global__ __launch_bounds__(kPassBlock) void kg_radial_heads(DevParams p, SampleArgs a, Records rt, int64_t nhead) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  if (i >= p.npix) return;
  uint32_t st = 0u;
  for (int64_t hd = blockIdx.y; hd < nhead; hd += gridDim.y)
    st |= radial_head_pixel<DO_PHI_T>(p.c, a.plan, p.npix, i, a.gflags[0], a.out, i, hd, rt);
  if (a.status && st) atomicOr(&a.status[i], st);
}

// Two waves per SIMD: the far pass needs about 268 VGPR+AGPR, one wave per
// SIMD unbounded; held to 256 with a few spills it runs about a quarter
// faster (0.32 against 0.44 ms at 512^2 x 500).
template <bool DO_PHI_T>
__global__ __launch_bounds__(kPassBlock) __attribute__((amdgpu_waves_per_eu(2))) void kg_radial_far(
    DevParams p, SampleArgs a, Records rt, int64_t nblk, FixupQueue qrep) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  if (i >= p.npix) return;
  const DevRepair repair{qrep};
  uint32_t st = 0u;
  for (int64_t b = blockIdx.y; b < nblk; b += gridDim.y)
    st |= radial_far_pixel<DO_PHI_T>(p.c, a.plan, p.npix, i, a.gflags[0], a.out, i, b, rt, repair);
  if (a.status && st) atomicOr(&a.status[i], st);
}

// one lane per (repair item, segment of its block) (grid-stride): each of
// a block's segment starts from the lost one on, by direct evaluation
constexpr int kRepairSegments = static_cast<int>(kAnchorBlock / kAnchorEvery);

template <bool DO_PHI_T>
__global__ __launch_bounds__(kPassBlock) void kg_radial_repair(DevParams p, SampleArgs a, Records rt,
                                                              FixupQueue qrep) {
  const unsigned long long c = *qrep.count;
  const unsigned long long n = (c < qrep.cap ? c : qrep.cap) * kRepairSegments;
  for (unsigned long long e = static_cast<unsigned long long>(blockIdx.x) * kPassBlock + threadIdx.x; e < n;
       e += static_cast<unsigned long long>(gridDim.x) * kPassBlock) {
    int64_t i, j;
    bool walked;
    unpack_item(qrep.items[e / kRepairSegments], i, j, walked);
    const int64_t sg = static_cast<int64_t>(e % kRepairSegments);
    const uint32_t st = radial_repair_pixel<DO_PHI_T>(p.c, a.plan, p.npix, i, a.gflags[0], a.out, i, j, sg, rt);
    if (a.status && st) atomicOr(&a.status[i], st);
  }
}

template <bool DO_PHI_T>
__global__ __launch_bounds__(kPassBlock) __attribute__((amdgpu_waves_per_eu(2))) void kg_radial_walk(DevParams p, SampleArgs a, Records rt, FixupQueue q) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPassBlock + threadIdx.x;
  if (i >= p.npix) return;
  const int64_t j0 = chunk_start(p.ngeo, blockIdx.y, a.chunks);
  const int64_t j1 = chunk_start(p.ngeo, blockIdx.y + 1, a.chunks);
  DevPush push{q};
  const uint32_t st = radial_walk_pixel<DO_PHI_T>(p.c, a.plan, p.npix, i, a.gflags[0], a.out, i, j0, j1, rt, push);
  if (a.status && st) atomicOr(&a.status[i], st);
}

template <bool DO_PHI_T>
// one lane per (queued range, sample of the range) (grid-stride)
__global__ __launch_bounds__(kPassBlock) void kg_radial_fixup(DevParams p, SampleArgs a, FixupQueue q) {
  const unsigned long long c = *q.count;
  const unsigned long long n = (c < q.cap ? c : q.cap) * kFixupSpan;
  for (unsigned long long e = static_cast<unsigned long long>(blockIdx.x) * kPassBlock + threadIdx.x; e < n;
       e += static_cast<unsigned long long>(gridDim.x) * kPassBlock) {
    int64_t i, j;
    bool walked;
    unpack_item(q.items[e / kFixupSpan], i, j, walked);
    const int64_t sg = static_cast<int64_t>(e % kFixupSpan);
    const uint32_t st =
        radial_fixup_sample<DO_PHI_T>(p.c, a.plan, p.npix, i, a.gflags[0], a.out, i, j, walked, sg);
    if (a.status && st) atomicOr(&a.status[i], st);
  }
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

// Per-sample counts (kerr_raytracing_utils.py:551-684): equatorial crossings
// and poloidal orbits, one lane per pixel (kg_counts.hpp).
__global__ __launch_bounds__(kBlock) void kg_counts(DevParams p, const double* __restrict__ alpha,
                                                    const double* __restrict__ beta, const double* __restrict__ tau,
                                                    int64_t ntau, double* __restrict__ n_eq,
                                                    double* __restrict__ n_pol) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const double be = beta[i];
  const CountConsts k = count_consts(p.c, alpha[i], be);
  for (int64_t j = 0; j < ntau; ++j) {
    const int64_t o = j * p.npix + i;
    const double tj = tau[o];
    if (n_eq) n_eq[o] = n_equatorial_at(k, p.north != 0, be, tj);
    if (n_pol) n_pol[o] = n_poloidal_at(k, tj);
  }
}

// Turning points passed by each Mino time (kerr_raytracing_utils.py:686-826),
// one lane per pixel: the reference's fractional n_angular_turnings, its
// floor (theta turning points) and the radial turning point (kg_counts.hpp).
__global__ __launch_bounds__(kBlock) void kg_turnings(DevParams p, const double* __restrict__ alpha,
                                                      const double* __restrict__ beta,
                                                      const double* __restrict__ tau, int64_t ntau,
                                                      double* __restrict__ n_ang, int32_t* __restrict__ turns_th,
                                                      int8_t* __restrict__ turns_r) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= p.npix) return;
  const double al = alpha[i], be = beta[i];
  const bool ang = n_ang || turns_th;
  CountConsts k{};
  if (ang) k = count_consts(p.c, al, be);
  RadialTurn rt{false, 0.0};
  if (turns_r) rt = radial_turn_consts(p.c, al, be);
  for (int64_t j = 0; j < ntau; ++j) {
    const int64_t o = j * p.npix + i;
    const double tj = tau[o];
    if (ang) {
      const double n = n_angular_at(k, be, tj);
      if (n_ang) n_ang[o] = n;
      // NaN (eta == 0 rays) maps to INT32_MIN, as numpy's float -> int32 cast does on x86
      if (turns_th) turns_th[o] = (n == n && kfabs(n) < 2.0e9) ? static_cast<int32_t>(kfloor(n)) : INT32_MIN;
    }
    if (turns_r) turns_r[o] = static_cast<int8_t>(radial_turns_at(rt, tj));
  }
}

// Samples per lane chunking: split ngeo so that about 2^21 (pixel, chunk)
// lanes exist (256 CUs x 128 waves of 64 over the kernels' lifetime: more
// lanes than the chip holds at once lets waves that finish early be
// replaced; measured at 512^2 x 500, 8 chunks run 2 % faster than 4 and as
// fast as 15), keeping chunks >= 32 samples so the walk's per-chunk set-up
// stays a few percent of the work.
int pick_chunks(int64_t npix, int64_t ngeo) {
  const int64_t target_lanes = 256LL * 128 * 64;
  int64_t c = (target_lanes + npix - 1) / (npix > 0 ? npix : 1);
  if (c < 1) c = 1;
  const int64_t max_c = ngeo / 32 > 0 ? ngeo / 32 : 1;
  if (c > max_c) c = max_c;
  if (c > 65535) c = 65535;
  return static_cast<int>(c);
}

int env_chunks() {
  const char* s = std::getenv("KGEO_CHUNKS");
  return s ? std::atoi(s) : 0;
}

// Workspace: the flag page, the plan [PL_N][npix] and, for a walked grid
// (ngeo > 2), the theta and radial record tables and the fixup queue (one
// item per kFixupSpan samples and pixel bounds every pass's pushes).
struct WsLayout {
  size_t plan, threc, rrec, queue, rqueue, total;
  int64_t nth, nr;
  unsigned long long qcap, rcap;
};

WsLayout ws_layout(int64_t npix, int64_t ngeo) {
  WsLayout L{};
  const size_t np = static_cast<size_t>(npix > 0 ? npix : 0);
  L.plan = kFlagBytes;
  size_t off = kFlagBytes + np * PL_N * sizeof(double);
  if (ngeo > 2) {
    L.nth = theta_records(ngeo);
    L.nr = radial_records(ngeo);
    L.threc = off;
    off += static_cast<size_t>(L.nth) * kThFields * np * sizeof(double);
    L.rrec = off;
    off += static_cast<size_t>(L.nr) * kRFields * np * sizeof(double);
    L.qcap = static_cast<unsigned long long>(np) * static_cast<unsigned long long>(fixup_items(ngeo));
    L.queue = off;
    off += static_cast<size_t>(L.qcap) * sizeof(unsigned long long);
    // the radial far pass's repair items, at most one per (pixel, block):
    // their own storage, as the radial anchor passes run beside the theta
    // passes, whose fixups use the queue above
    L.rcap = static_cast<unsigned long long>(np) * static_cast<unsigned long long>(anchor_blocks(ngeo));
    L.rqueue = off;
    off += static_cast<size_t>(L.rcap) * sizeof(unsigned long long);
  }
  L.total = off;
  return L;
}

// The kernels always sample t, phi and sigma (as the reference's
// raytrace_ana does, kerr_raytracing_ana.py:97): the theta pass parks G_phi,
// G_t in the t and phi outputs and the plan holds the phi/t constants.  A
// request without KGEO_DO_PHI_T is refused rather than run on kernel code
// that no test has run.
int require_phi_t(const kgeo_ray_params* p) {
  if (!(p->flags & KGEO_DO_PHI_T))
    return fail_msg(KGEO_E_ARG, "KGEO_DO_PHI_T is required (t, phi and sigma are always sampled)");
  return KGEO_OK;
}

// Sizes one call handles: a fixup item packs the pixel in 32 bits and the
// sample in 31, and the queue counters are 32-bit (a pass queues at most
// npix * ceil(ngeo / 8) items).  Larger jobs are split over pixels.
int size_limit(int64_t npix, int64_t ngeo) {
  if (npix > 0x7fffffffLL) return fail(KGEO_E_ARG, "npix above 2^31 - 1%s (%lld): split the pixels", "", npix);
  if (ngeo > 0x7fffffffLL) return fail(KGEO_E_ARG, "ngeo above 2^31 - 1%s (%lld)", "", ngeo);
  if (npix > 0 && ngeo > 2 && npix * fixup_items(ngeo) > 0xffffffffLL)
    return fail(KGEO_E_ARG, "npix * ceil(ngeo / 8) above 2^32%s (%lld): split the pixels", "",
                npix * fixup_items(ngeo));
  return KGEO_OK;
}

int validate(const kgeo_ray_params* p, const double* alpha, const double* beta, void* ws, size_t wsb,
             int64_t ngeo) {
  if (!p) return fail_msg(KGEO_E_ARG, "params is NULL");
  if (require_phi_t(p)) return KGEO_E_ARG;
  if (p->npix < 0) return fail(KGEO_E_ARG, "npix < 0%s (%lld)", "", p->npix);
  if (size_limit(p->npix, ngeo)) return KGEO_E_ARG;
  if (p->npix > 0 && (!alpha || !beta)) return fail_msg(KGEO_E_ARG, "alpha/beta is NULL");
  if (p->npix > 0 && !ws) return fail_msg(KGEO_E_WORKSPACE, "workspace is NULL");
  const size_t need = kgeo_workspace_bytes(p->npix, ngeo);
  if (wsb < need)
    return fail(KGEO_E_WORKSPACE, "workspace too small%s: need %lld bytes", "", static_cast<long long>(need));
  if (reinterpret_cast<uintptr_t>(ws) % 256 != 0) return fail_msg(KGEO_E_WORKSPACE, "workspace not 256-byte aligned");
  return KGEO_OK;
}


// A second stream per device for the radial anchor passes, which depend
// only on the plan and so run beside the theta passes (two latency-bound
// kernels interleave on the SIMDs); created once per device, with the fork
// and join events.  The fork, the side stream's work and the join are
// enqueued under one lock, so calls from several host threads keep their
// own dependencies (their radial anchor passes take turns on the side
// stream).  Streams keep the default priority.
struct SideStream {
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t planned = nullptr;  // the side stream's radial plan constants are written
  bool tried = false;
};
constexpr int kMaxDev = 64;
std::mutex g_side_mu;
SideStream g_side[kMaxDev];

// The side stream of the current device, or nullptr (the passes then run in
// order on the caller's stream): also when the caller's stream belongs to
// another device than the current one, whose side stream it cannot fork to.
// Call with g_side_mu held.
SideStream* side_stream(hipStream_t st) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return nullptr;
  if (st) {
    hipDevice_t sd = -1;
    if (hipStreamGetDevice(st, &sd) != hipSuccess || static_cast<int>(sd) != dev) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  SideStream& s = g_side[dev];
  if (!s.tried) {
    s.tried = true;
    if (hipStreamCreateWithFlags(&s.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.planned, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      s.side = nullptr;
    }
  }
  return s.side ? &s : nullptr;
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
  const unsigned nb = static_cast<unsigned>((d.npix + kPassBlock - 1) / kPassBlock);
  std::lock_guard<std::mutex> lock(g_side_mu);
  SideStream* ss = side_stream(st);
  if (!ss) {
    hipLaunchKernelGGL(kg_setup, dim3(nb), dim3(kPassBlock), 0, st, d, alpha, beta, plan, gflags, rclass, thclass,
                       nmax_eq, tau_tot, status);
    return check_launch("kg_setup");
  }
  // the radial constants on the side stream, after everything enqueued so
  // far on the caller's (an earlier sample pass may still read the plan);
  // kgeo_sample's radial passes follow them there, and its direct path
  // waits for `planned`
  if (hipEventRecord(ss->fork, st) != hipSuccess || hipStreamWaitEvent(ss->side, ss->fork, 0) != hipSuccess)
    return fail_msg(KGEO_E_HIP, "fork to the side stream failed");
  hipLaunchKernelGGL(kg_setup_theta, dim3(nb), dim3(kPassBlock), 0, st, d, alpha, beta, plan, gflags, rclass, thclass,
                     nmax_eq, tau_tot, status);
  if ((rc = check_launch("kg_setup_theta"))) return rc;
  hipLaunchKernelGGL(kg_setup_radial, dim3(nb), dim3(kPassBlock), 0, ss->side, d, alpha, beta, plan);
  if ((rc = check_launch("kg_setup_radial"))) return rc;
  if (hipEventRecord(ss->planned, ss->side) != hipSuccess) return fail_msg(KGEO_E_HIP, "plan event failed");
  return KGEO_OK;
}

template <bool DO_PHI_T>
int run_walk(const DevParams& d, const SampleArgs& a, void* ws, dim3 grid, hipStream_t st) {
  const WsLayout L = ws_layout(d.npix, d.ngeo);
  char* base = static_cast<char*>(ws);
  uint32_t* flags = static_cast<uint32_t*>(ws);
  const Records rth{reinterpret_cast<double*>(base + L.threc), d.npix, 0};
  const Records rr{reinterpret_cast<double*>(base + L.rrec), d.npix, 0};
  unsigned long long* items = reinterpret_cast<unsigned long long*>(base + L.queue);
  unsigned long long* ritems = reinterpret_cast<unsigned long long*>(base + L.rqueue);
  const FixupQueue qth{items, flags + kThetaCount, L.qcap};
  const FixupQueue qr{items, flags + kRadialCount, L.qcap};
  const FixupQueue qrep{ritems, flags + kRepairCount, L.rcap};
  const unsigned nb = static_cast<unsigned>((d.npix + kPassBlock - 1) / kPassBlock);
  const dim3 wgrid(nb, grid.y);
  const unsigned long long fb = (L.qcap + kPassBlock - 1) / kPassBlock;
  const dim3 fgrid(static_cast<unsigned>(fb < 4096ull ? fb : 4096ull));
  const unsigned long long rb = (L.rcap * kRepairSegments + kPassBlock - 1) / kPassBlock;
  const dim3 rgrid(static_cast<unsigned>(rb < 4096ull ? (rb > 0 ? rb : 1ull) : 4096ull));
  auto recs = [](int64_t n) { return static_cast<unsigned>(n < 65535 ? n : 65535); };
  hipLaunchKernelGGL(kg_clear_counters, dim3(1), dim3(64), 0, st, flags);
  int rc = check_launch("kg_clear_counters");
  if (rc) return rc;
  const int64_t nblk = anchor_blocks(d.ngeo), nhead = anchor_heads(d.ngeo);
  {
    std::lock_guard<std::mutex> lock(g_side_mu);
    SideStream* ss = side_stream(st);
    const hipStream_t rs = ss ? ss->side : st;
    if (ss) {
      if (hipEventRecord(ss->fork, st) != hipSuccess || hipStreamWaitEvent(rs, ss->fork, 0) != hipSuccess)
        return fail_msg(KGEO_E_HIP, "fork to the side stream failed");
    }
    // radial heads, far and repair: records only (the walk emits the samples)
    hipLaunchKernelGGL(kg_radial_heads<DO_PHI_T>, dim3(nb, recs(nhead)), dim3(kPassBlock), 0, rs, d, a, rr, nhead);
    if ((rc = check_launch("kg_radial_heads"))) return rc;
    hipLaunchKernelGGL(kg_radial_far<DO_PHI_T>, dim3(nb, recs(nblk)), dim3(kPassBlock), 0, rs, d, a, rr, nblk,
                       qrep);
    if ((rc = check_launch("kg_radial_far"))) return rc;
    hipLaunchKernelGGL(kg_radial_repair<DO_PHI_T>, rgrid, dim3(kPassBlock), 0, rs, d, a, rr, qrep);
    if ((rc = check_launch("kg_radial_repair"))) return rc;
    // theta's last-sample head: read by the radial walk only
    hipLaunchKernelGGL(kg_theta_heads<DO_PHI_T>, dim3(nb, 1), dim3(kPassBlock), 0, rs, d, a, rth, nhead, 1);
    if ((rc = check_launch("kg_theta_heads (last sample)"))) return rc;
    if (ss && hipEventRecord(ss->join, rs) != hipSuccess) return fail_msg(KGEO_E_HIP, "join event failed");
    // theta passes on the caller's stream meanwhile
    hipLaunchKernelGGL(kg_theta_heads<DO_PHI_T>, dim3(nb, recs(nhead - 1)), dim3(kPassBlock), 0, st, d, a, rth, nhead,
                       0);
    if ((rc = check_launch("kg_theta_heads"))) return rc;
    hipLaunchKernelGGL(kg_theta_far<DO_PHI_T>, dim3(nb, recs(nblk)), dim3(kPassBlock), 0, st, d, a, rth, nblk);
    if ((rc = check_launch("kg_theta_far"))) return rc;
    hipLaunchKernelGGL(kg_theta_walk<DO_PHI_T>, grid, dim3(kBlock), 0, st, d, a, rth, qth);
    if ((rc = check_launch("kg_theta_walk"))) return rc;
    hipLaunchKernelGGL(kg_theta_fixup<DO_PHI_T>, fgrid, dim3(kPassBlock), 0, st, d, a, qth);
    if ((rc = check_launch("kg_theta_fixup"))) return rc;
    if (ss && hipStreamWaitEvent(st, ss->join, 0) != hipSuccess) return fail_msg(KGEO_E_HIP, "join failed");
  }
  // the radial walk emits every radial sample it walks and every segment
  // start from its record, combining with the theta pass's G_phi, G_t
  hipLaunchKernelGGL(kg_radial_walk<DO_PHI_T>, wgrid, dim3(kPassBlock), 0, st, d, a, rr, qr);
  if ((rc = check_launch("kg_radial_walk"))) return rc;
  hipLaunchKernelGGL(kg_radial_fixup<DO_PHI_T>, fgrid, dim3(kPassBlock), 0, st, d, a, qr);
  return check_launch("kg_radial_fixup");
}

int run_sample(const kgeo_ray_params* p, int64_t ngeo, const double* tau_in, double* tau, double* t, double* r,
               double* th, double* ph, double* sig, uint32_t* status, void* ws, hipStream_t st) {
  DevParams d = make_dev(p, ngeo);
  if (d.npix == 0 || ngeo <= 0) return KGEO_OK;
  if (!t || !ph) return fail_msg(KGEO_E_ARG, "t and ph outputs are required");
  const uint32_t* gflags = static_cast<const uint32_t*>(ws);
  const double* plan = reinterpret_cast<const double*>(static_cast<const char*>(ws) + kFlagBytes);
  const unsigned nb = static_cast<unsigned>((d.npix + kBlock - 1) / kBlock);
  const int chunks = kgeo_sample_chunks(d.npix, ngeo);
  dim3 grid(nb, static_cast<unsigned>(chunks));
  SampleArgs a{plan, gflags, SampleOut{tau_in, tau, t, r, th, ph, sig, d.npix, ngeo}, status, chunks};
  // the linspace grid walks (heads, far, walk, fixup passes); caller-given
  // Mino times and grids of at most two samples are evaluated directly
  if (!tau_in && ngeo > 2) return run_walk<true>(d, a, ws, grid, st);
  {
    // the plan's radial constants come from the side stream (run_plan)
    std::lock_guard<std::mutex> lock(g_side_mu);
    SideStream* ss = side_stream(st);
    if (ss && hipStreamWaitEvent(st, ss->planned, 0) != hipSuccess)
      return fail_msg(KGEO_E_HIP, "wait for the plan failed");
  }
  hipLaunchKernelGGL(kg_theta<true>, grid, dim3(kBlock), 0, st, d, a);
  int rc = check_launch("kg_theta");
  if (rc) return rc;
  hipLaunchKernelGGL(kg_radial<true>, grid, dim3(kBlock), 0, st, d, a);
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

int kgeo_sample_chunks(int64_t npix, int64_t ngeo) {
  if (npix <= 0 || ngeo <= 0) return 0;
  int c = env_chunks();
  if (c <= 0) c = pick_chunks(npix, ngeo);
  if (c > ngeo) c = static_cast<int>(ngeo);
  return c;
}

size_t kgeo_workspace_bytes(int64_t npix, int64_t ngeo) {
  if (npix < 0) return 0;
  return ws_layout(npix, ngeo).total;
}

int kgeo_plan(const kgeo_ray_params* p, const double* alpha, const double* beta, int8_t* rclass, int8_t* thclass,
              int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace, size_t workspace_bytes,
              void* stream) {
  int rc = validate(p, alpha, beta, workspace, workspace_bytes, 0);
  if (rc) return rc;
  return run_plan(p, alpha, beta, rclass, thclass, nmax_eq, tau_tot, status, workspace,
                  static_cast<hipStream_t>(stream));
}

int kgeo_sample(const kgeo_ray_params* p, const double* tau_in, int64_t ntau, double* tau, double* t, double* r,
                double* th, double* ph, double* sig, uint32_t* status, void* workspace, size_t workspace_bytes,
                void* stream) {
  if (!p) return fail_msg(KGEO_E_ARG, "params is NULL");
  if (require_phi_t(p)) return KGEO_E_ARG;
  if (p->npix < 0 || ntau < 0) return fail_msg(KGEO_E_ARG, "negative size");
  if (size_limit(p->npix, tau_in ? 0 : ntau)) return KGEO_E_ARG;
  if (p->npix > 0 && !workspace) return fail_msg(KGEO_E_WORKSPACE, "workspace is NULL");
  const size_t need = kgeo_workspace_bytes(p->npix, tau_in ? 0 : ntau);
  if (workspace_bytes < need)
    return fail(KGEO_E_WORKSPACE, "workspace too small%s: need %lld bytes", "", static_cast<long long>(need));
  return run_sample(p, ntau, tau_in, tau, t, r, th, ph, sig, status, workspace, static_cast<hipStream_t>(stream));
}

int kgeo_raytrace_ana(const kgeo_ray_params* p, const double* alpha, const double* beta, double* tau, double* t,
                      double* r, double* th, double* ph, double* sig, int8_t* rclass, int8_t* thclass,
                      int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace, size_t workspace_bytes,
                      void* stream) {
  int rc = validate(p, alpha, beta, workspace, workspace_bytes, p ? p->ngeo : 0);
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
  int rc = validate(p, alpha, beta, workspace, workspace_bytes, 0);
  if (rc) return rc;
  if (ntau < 0) return fail(KGEO_E_ARG, "ntau < 0%s (%lld)", "", ntau);
  if (ntau > 0 && p->npix > 0 && !tau_in) return fail_msg(KGEO_E_ARG, "tau_in is NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  rc = run_plan(p, alpha, beta, nullptr, nullptr, nullptr, nullptr, status, workspace, st);
  if (rc) return rc;
  return run_sample(p, ntau, tau_in, nullptr, t, r, th, ph, sig, status, workspace, st);
}

int kgeo_crossings(const kgeo_ray_params* p, const double* alpha, const double* beta, const double* tau,
                   int64_t ntau, double* n_equatorial, double* n_poloidal, void* stream) {
  if (!p) return fail_msg(KGEO_E_ARG, "params is NULL");
  if (p->npix < 0 || ntau < 0) return fail_msg(KGEO_E_ARG, "negative size");
  if (p->npix == 0 || ntau == 0) return KGEO_OK;
  if (!alpha || !beta || !tau) return fail_msg(KGEO_E_ARG, "NULL array");
  DevParams d = make_dev(p, ntau);
  const unsigned nb = static_cast<unsigned>((d.npix + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_counts, dim3(nb), dim3(kBlock), 0, static_cast<hipStream_t>(stream), d, alpha, beta, tau,
                     ntau, n_equatorial, n_poloidal);
  return check_launch("kg_counts");
}

int kgeo_turnings(const kgeo_ray_params* p, const double* alpha, const double* beta, const double* tau,
                  int64_t ntau, double* n_angular, int32_t* turns_theta, int8_t* turns_r, void* stream) {
  if (!p) return fail_msg(KGEO_E_ARG, "params is NULL");
  if (p->npix < 0 || ntau < 0) return fail_msg(KGEO_E_ARG, "negative size");
  // n_angular_turnings accepts th_o in [0, pi/2] only (kerr_raytracing_utils.py:690-691)
  if ((n_angular || turns_theta) && !(p->th_o >= 0.0 && p->th_o <= kHalfPi))
    return fail_msg(KGEO_E_ARG, "th_o should be a float in range (0,pi/2]");
  if (p->npix == 0 || ntau == 0) return KGEO_OK;
  if (!alpha || !beta || !tau) return fail_msg(KGEO_E_ARG, "NULL array");
  DevParams d = make_dev(p, ntau);
  const unsigned nb = static_cast<unsigned>((d.npix + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_turnings, dim3(nb), dim3(kBlock), 0, static_cast<hipStream_t>(stream), d, alpha, beta, tau,
                     ntau, n_angular, turns_theta, turns_r);
  return check_launch("kg_turnings");
}

int kgeo_ellip_pi(const double* n, const double* phi, const double* m, double* out, int64_t count, void* stream) {
  if (count < 0) return fail(KGEO_E_ARG, "count < 0%s (%lld)", "", count);
  if (count == 0) return KGEO_OK;
  if (!n || !phi || !m || !out) return fail_msg(KGEO_E_ARG, "NULL array");
  const unsigned nb = static_cast<unsigned>((count + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kg_ellip_pi, dim3(nb), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, phi, m, out, count);
  return check_launch("kg_ellip_pi");
}

}  // extern "C"
