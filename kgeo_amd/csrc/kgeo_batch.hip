// Batched survey: many raytrace_ana images, each with its own spin,
// inclination, resolution and ngeo, in one persistent launch that is
// captured once in a HIP graph and replayed.
//
//   kg_batch_clear       zero the per-image flag words, the per-image
//                        completed-item counters and the work counter
//   kg_batch_setup       one lane per pixel slot of every image: the same
//                        per-pixel plan and labels as kg_setup.  Each image's
//                        slots are padded to a multiple of 64, so a wave never
//                        straddles two images and the image-wide predicates
//                        (any eta > 0, any zero grid step) are per-image
//                        wave ballots.
//   kg_batch_persistent  a fixed grid of waves; each wave takes work items
//                        (image, 64-pixel tile, Mino-time chunk) from one
//                        atomic counter until the queue is empty and runs the
//                        theta loop then the radial loop of that item
//                        (kg_sample.hpp, the code the plain kernels run).
//                        Every wave leaves once the counter passes the item
//                        total, so the grid always drains.
// Outputs go to caller arenas: image k's [ngeo_k][npix_k] block of r, theta,
// phi, t starts at element out_off[k] (sum of npix*ngeo of the images before
// it); per-pixel inputs and labels are concatenated at pix_off[k].

#include <hip/hip_runtime.h>

#include <cstdio>
#include <new>
#include <vector>

#include "../../include/kgeo.h"
#include "kg_sample.hpp"

using namespace kg;

namespace {

thread_local char g_err_b[512] = "";

int bfail(int code, const char* msg, long long v = 0) {
  std::snprintf(g_err_b, sizeof(g_err_b), msg, v);
  return code;
}

int blaunch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::snprintf(g_err_b, sizeof(g_err_b), "%s failed: %d", what, static_cast<int>(e));
    return KGEO_E_HIP;
  }
  return KGEO_OK;
}

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int64_t kMaxItems = (1LL << 31) - (1LL << 24);  // leaves room for one exit fetch per wave

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

struct BatchImage {
  RayScalars<double> c;
  int64_t npix, ngeo;
  int64_t pix_off;   // first pixel in the concatenated per-pixel arrays
  int64_t slot_off;  // first plan column (64-aligned)
  int64_t out_off;   // first element of the image's [ngeo][npix] block
  int64_t item_off;  // first work item
  int32_t chunks;    // Mino-time chunks per pixel
  int32_t do_phi_t;
};

struct BatchDev {
  const BatchImage* imgs;
  int nimg;
  int64_t total_slots;
  uint32_t total_items;
  const double* alpha;
  const double* beta;
  double* plan;       // [PL_N][total_slots]
  uint32_t* flags;    // [nimg]
  uint32_t* done;     // [nimg] completed items
  uint32_t* counter;
  double *t, *r, *th, *ph;
  int8_t* rclass;
  int8_t* thclass;
  int32_t* nmax_eq;
  double* tau_tot;
  uint32_t* status;
};

// last image k with key(k) <= key; every lane of a wave searches the same key
__device__ inline int find_image(const BatchImage* imgs, int n, int64_t key, bool by_item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t off = by_item ? imgs[mid].item_off : imgs[mid].slot_off;
    if (off <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void kg_batch_clear(BatchDev d) {
  for (int k = threadIdx.x; k < d.nimg; k += blockDim.x) {
    d.flags[k] = 0u;
    d.done[k] = 0u;
  }
  if (threadIdx.x == 0) d.counter[0] = 0u;
}

__global__ __launch_bounds__(kBlock) void kg_batch_setup(BatchDev d) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  // total_slots is a multiple of 64: whole waves leave here together
  if (s >= d.total_slots) return;
  const int im = find_image(d.imgs, d.nimg, s, false);
  const BatchImage& B = d.imgs[im];
  const int64_t i = s - B.slot_off;
  SetupOut so{false, false};
  if (i < B.npix) {
    const int64_t g = B.pix_off + i;
    so = setup_pixel(B.c, B.ngeo, B.do_phi_t != 0, d.alpha[g], d.beta[g], d.plan, d.total_slots, s, d.rclass,
                     d.thclass, d.nmax_eq, d.tau_tot, d.status, g);
  }
  wave_or(&d.flags[im], so.eta_pos, GF_ANY_ETA_POS);
  wave_or(&d.flags[im], so.step_zero, GF_ANY_STEP_ZERO);
}

// The work-item fetch is written so that the compiler sees it as uniform:
// every lane issues the atomic (lane 0 adds 1, the others 0; the backend
// combines them into one atomic per wave) and the value is taken with
// readfirstlane, so `item` lives in a scalar register and the exit test is
// a scalar branch.  A lane-0-only atomic followed by a lane shuffle lets the
// structurizer split the wave's lanes into two loop nests, in which the
// shuffle re-reads an inactive lane 0 and the loop never ends.
__device__ inline uint32_t fetch_item(uint32_t* counter, int lane) {
  const uint32_t got = atomicAdd(counter, lane == 0 ? 1u : 0u);
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(got)));
}

__global__ __launch_bounds__(kBlock) void kg_batch_persistent(BatchDev d) {
  const int lane = __lane_id();
  for (;;) {
    const uint32_t item = fetch_item(d.counter, lane);
    if (item >= d.total_items) break;  // scalar branch: the whole wave leaves once the queue is empty
    const int im = find_image(d.imgs, d.nimg, static_cast<int64_t>(item), true);
    const BatchImage& B = d.imgs[im];
    const int64_t li = static_cast<int64_t>(item) - B.item_off;
    const int64_t tile = li / B.chunks;
    const int64_t chunk = li % B.chunks;
    const int64_t i = tile * kWave + lane;
    if (i < B.npix) {
      const int64_t s = B.slot_off + i;
      const int64_t j0 = chunk_start(B.ngeo, chunk, B.chunks);
      const int64_t j1 = chunk_start(B.ngeo, chunk + 1, B.chunks);
      const uint32_t fl = d.flags[im];
      const bool phi_t = B.do_phi_t != 0;
      SampleOut o{nullptr, nullptr, phi_t ? d.t + B.out_off : nullptr, d.r + B.out_off, d.th + B.out_off,
                  phi_t ? d.ph + B.out_off : nullptr, nullptr, B.npix, B.ngeo};
      uint32_t st;
      if (phi_t) {
        st = theta_pixel<true>(d.plan, d.total_slots, s, fl, o, i, j0, j1);
        st |= radial_pixel<true>(B.c, d.plan, d.total_slots, s, fl, o, i, j0, j1);
      } else {
        st = theta_pixel<false>(d.plan, d.total_slots, s, fl, o, i, j0, j1);
        st |= radial_pixel<false>(B.c, d.plan, d.total_slots, s, fl, o, i, j0, j1);
      }
      if (d.status && st) atomicOr(&d.status[B.pix_off + i], st);
    }
    atomicAdd(&d.done[im], lane == 0 ? 1u : 0u);
  }
}

RayScalars<double> scalars_of(const kgeo_ray_params& p) {
  RayScalars<double> c = make_scalars(p.a, p.th_o, p.r_o, p.maxtaufrac);
  c.a2 = p.a2;
  c.sin_th = p.sin_th;
  c.cos_th = p.cos_th;
  c.cos2_th = p.cos2_th;
  c.rplus = p.rplus;
  c.rminus = p.rminus;
  return c;
}

int check_images(const kgeo_ray_params* imgs, int32_t nimg) {
  if (!imgs || nimg <= 0) return bfail(KGEO_E_ARG, "no images (nimg = %lld)", nimg);
  for (int32_t k = 0; k < nimg; ++k) {
    if (imgs[k].npix < 1) return bfail(KGEO_E_ARG, "image %lld has npix < 1", k);
    if (imgs[k].ngeo < 1) return bfail(KGEO_E_ARG, "image %lld has ngeo < 1", k);
  }
  return KGEO_OK;
}

int64_t total_slots_of(const kgeo_ray_params* imgs, int32_t nimg) {
  int64_t s = 0;
  for (int32_t k = 0; k < nimg; ++k) s += (imgs[k].npix + kWave - 1) / kWave * kWave;
  return s;
}

}  // namespace

struct kgeo_batch {
  BatchDev d;
  std::vector<uint32_t> items_per_image;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" {

const char* kgeo_batch_last_error(void) { return g_err_b; }

size_t kgeo_batch_workspace_bytes(const kgeo_ray_params* imgs, int32_t nimg) {
  if (check_images(imgs, nimg) != KGEO_OK) return 0;
  return align256(sizeof(BatchImage) * static_cast<size_t>(nimg)) +
         2 * align256(sizeof(uint32_t) * static_cast<size_t>(nimg)) + 256 +
         static_cast<size_t>(total_slots_of(imgs, nimg)) * PL_N * sizeof(double);
}

int kgeo_batch_layout(const kgeo_ray_params* imgs, int32_t nimg, int64_t* pix_off, int64_t* out_off) {
  int rc = check_images(imgs, nimg);
  if (rc) return rc;
  int64_t p = 0, o = 0;
  for (int32_t k = 0; k < nimg; ++k) {
    if (pix_off) pix_off[k] = p;
    if (out_off) out_off[k] = o;
    p += imgs[k].npix;
    o += imgs[k].npix * imgs[k].ngeo;
  }
  if (pix_off) pix_off[nimg] = p;
  if (out_off) out_off[nimg] = o;
  return KGEO_OK;
}

int kgeo_batch_plan(const kgeo_ray_params* imgs, int32_t nimg, const kgeo_batch_tuning* tune, const double* alpha,
                    const double* beta, double* t, double* r, double* th, double* ph, int8_t* rclass,
                    int8_t* thclass, int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace,
                    size_t workspace_bytes, void* stream, kgeo_batch** out) {
  if (!out) return bfail(KGEO_E_ARG, "out handle is NULL");
  *out = nullptr;
  int rc = check_images(imgs, nimg);
  if (rc) return rc;
  if (!alpha || !beta || !th || !r) return bfail(KGEO_E_ARG, "alpha, beta, th and r are required");
  if (!stream) return bfail(KGEO_E_ARG, "graph capture needs a non-default stream");
  const size_t need = kgeo_batch_workspace_bytes(imgs, nimg);
  if (!workspace || workspace_bytes < need)
    return bfail(KGEO_E_WORKSPACE, "workspace too small: need %lld bytes", static_cast<long long>(need));
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0) return bfail(KGEO_E_WORKSPACE, "workspace not 256-byte aligned");
  kgeo_batch_tuning tn{64, 8, {0, 0}};
  if (tune) tn = *tune;
  if (tn.samples_per_item < 1) return bfail(KGEO_E_ARG, "samples_per_item < 1 (%lld)", tn.samples_per_item);
  if (tn.waves_per_cu < 1 || tn.waves_per_cu > 64)
    return bfail(KGEO_E_ARG, "waves_per_cu outside 1..64 (%lld)", tn.waves_per_cu);

  std::vector<BatchImage> host(nimg);
  std::vector<uint32_t> per(nimg);
  int64_t pix = 0, slots = 0, outo = 0, items = 0;
  bool any_phi = false;
  for (int32_t k = 0; k < nimg; ++k) {
    const kgeo_ray_params& p = imgs[k];
    BatchImage& B = host[k];
    B.c = scalars_of(p);
    B.npix = p.npix;
    B.ngeo = p.ngeo;
    B.pix_off = pix;
    B.slot_off = slots;
    B.out_off = outo;
    B.item_off = items;
    // items hold whole anchor intervals (chunk starts are multiples of kAnchorEveryLong)
    const int64_t spi =
        (static_cast<int64_t>(tn.samples_per_item) + kAnchorEveryLong - 1) / kAnchorEveryLong * kAnchorEveryLong;
    const int64_t ch = (p.ngeo + spi - 1) / spi;
    B.chunks = static_cast<int32_t>(ch);
    B.do_phi_t = (p.flags & KGEO_DO_PHI_T) ? 1 : 0;
    any_phi |= B.do_phi_t != 0;
    const int64_t tiles = (p.npix + kWave - 1) / kWave;
    pix += p.npix;
    slots += tiles * kWave;
    outo += p.npix * p.ngeo;
    items += tiles * ch;
    if (items > kMaxItems) return bfail(KGEO_E_ARG, "too many work items (%lld); raise samples_per_item", items);
    per[k] = static_cast<uint32_t>(tiles * ch);
  }
  if (any_phi && (!t || !ph)) return bfail(KGEO_E_ARG, "t and ph are required with KGEO_DO_PHI_T");

  char* w = static_cast<char*>(workspace);
  BatchDev d;
  d.imgs = reinterpret_cast<const BatchImage*>(w);
  w += align256(sizeof(BatchImage) * static_cast<size_t>(nimg));
  d.flags = reinterpret_cast<uint32_t*>(w);
  w += align256(sizeof(uint32_t) * static_cast<size_t>(nimg));
  d.done = reinterpret_cast<uint32_t*>(w);
  w += align256(sizeof(uint32_t) * static_cast<size_t>(nimg));
  d.counter = reinterpret_cast<uint32_t*>(w);
  w += 256;
  d.plan = reinterpret_cast<double*>(w);
  d.nimg = nimg;
  d.total_slots = slots;
  d.total_items = static_cast<uint32_t>(items);
  d.alpha = alpha;
  d.beta = beta;
  d.t = t;
  d.r = r;
  d.th = th;
  d.ph = ph;
  d.rclass = rclass;
  d.thclass = thclass;
  d.nmax_eq = nmax_eq;
  d.tau_tot = tau_tot;
  d.status = status;

  hipStream_t st = static_cast<hipStream_t>(stream);
  // the image descriptors are uploaded once, outside the graph
  if (hipMemcpyAsync(const_cast<BatchImage*>(d.imgs), host.data(), sizeof(BatchImage) * host.size(),
                     hipMemcpyHostToDevice, st) != hipSuccess)
    return bfail(KGEO_E_HIP, "descriptor upload failed");
  if (hipStreamSynchronize(st) != hipSuccess) return bfail(KGEO_E_HIP, "descriptor upload failed (sync)");
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return bfail(KGEO_E_HIP, "hipGetDevice failed");
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return bfail(KGEO_E_HIP, "CU count query failed");
  const int waves_per_block = kBlock / kWave;
  const unsigned blocks = static_cast<unsigned>((cus * tn.waves_per_cu + waves_per_block - 1) / waves_per_block);
  const unsigned setup_blocks = static_cast<unsigned>((slots + kBlock - 1) / kBlock);

  kgeo_batch* b = new (std::nothrow) kgeo_batch;
  if (!b) return bfail(KGEO_E_ARG, "out of host memory");
  b->d = d;
  b->items_per_image = per;
  b->graph = nullptr;
  b->exec = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    delete b;
    return bfail(KGEO_E_HIP, "hipStreamBeginCapture failed");
  }
  hipLaunchKernelGGL(kg_batch_clear, dim3(1), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(kg_batch_setup, dim3(setup_blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(kg_batch_persistent, dim3(blocks), dim3(kBlock), 0, st, d);
  const int lrc = blaunch("batch kernel launch (capture)");
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(st, &g);
  if (lrc || ce != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    delete b;
    return lrc ? lrc : bfail(KGEO_E_HIP, "hipStreamEndCapture failed");
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    delete b;
    return bfail(KGEO_E_HIP, "hipGraphInstantiate failed");
  }
  *out = b;
  return KGEO_OK;
}

int kgeo_batch_run(kgeo_batch* b, void* stream) {
  if (!b) return bfail(KGEO_E_ARG, "batch handle is NULL");
  if (hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)) != hipSuccess)
    return bfail(KGEO_E_HIP, "hipGraphLaunch failed");
  return KGEO_OK;
}

int64_t kgeo_batch_items(const kgeo_batch* b) { return b ? static_cast<int64_t>(b->d.total_items) : -1; }

int kgeo_batch_progress(const kgeo_batch* b, uint32_t* done_items, uint32_t* total_items) {
  if (!b) return bfail(KGEO_E_ARG, "batch handle is NULL");
  const size_t n = static_cast<size_t>(b->d.nimg);
  if (total_items)
    for (size_t k = 0; k < n; ++k) total_items[k] = b->items_per_image[k];
  if (done_items && hipMemcpy(done_items, b->d.done, n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
    return bfail(KGEO_E_HIP, "progress copy failed");
  return KGEO_OK;
}

int kgeo_batch_destroy(kgeo_batch* b) {
  if (!b) return KGEO_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  delete b;
  return KGEO_OK;
}

}  // extern "C"
