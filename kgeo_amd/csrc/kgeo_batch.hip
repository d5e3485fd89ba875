// Batched survey: many raytrace_ana images, each with its own spin,
// inclination, resolution and ngeo, in persistent launches that are
// captured once in a HIP graph and replayed.
//
//   kg_batch_clear       zero the per-image flag words and item counters,
//                        the pass counters and the fixup-queue counters
//   kg_batch_setup       one lane per pixel slot of every image: the same
//                        per-pixel plan and labels as kg_setup.  Each image's
//                        slots are padded to a multiple of 64, so a wave never
//                        straddles two images and the image-wide predicates
//                        (any eta > 0, any zero grid step) are per-image
//                        wave ballots.
//   kg_batch_pass<P>     for theta then r, the heads, far and walk passes of
//                        the plain kernels (kg_sample.hpp), each a fixed grid
//                        of waves taking work items (image, 64-pixel tile,
//                        head / block / Mino-time chunk) from its own atomic
//                        counter until the pass is empty; every wave leaves
//                        once the counter passes the pass's item total, so
//                        the grid always drains;
//   kg_batch_repair      after the radial far pass: the blocks whose long
//                        step was refused, one lane per (block, segment),
//                        by direct evaluation (grid-stride);
//   kg_batch_fixup<R>    the queued ranges of all images, one lane per
//                        (range, sample), by direct evaluation (grid-stride).
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

// pass kinds and their per-tile work units
enum : int { K_HEAD = 0, K_FAR = 1, K_WALK = 2, K_N = 3 };
// pass counters: theta heads, far, walk, then r heads, far, walk; then the
// theta and radial fixup-queue counters
constexpr int kPassCounters = 2 * K_N;
constexpr int kQueueCounter = kPassCounters;  // + 0 theta, + 1 r
constexpr int kRepairCounter = kPassCounters + 2;  // the radial far pass's repair items
constexpr int kCounters = kPassCounters + 3;

struct BatchImage {
  RayScalars<double> c;
  int64_t npix, ngeo;
  int64_t pix_off;   // first pixel in the concatenated per-pixel arrays
  int64_t slot_off;  // first plan column (64-aligned)
  int64_t slots;     // padded pixel count
  int64_t out_off;   // first element of the image's [ngeo][npix] block
  int64_t item_off[K_N];  // first work item of each pass kind
  int64_t th_rec, r_rec;  // first element of the image's record tables
  int32_t units[K_N];     // heads, anchor blocks, Mino-time chunks per tile
  int32_t do_phi_t;
};

struct BatchDev {
  const BatchImage* imgs;
  int nimg;
  int64_t total_slots;
  uint32_t total_items[K_N];
  const double* alpha;
  const double* beta;
  double* plan;       // [PL_N][total_slots]
  double* rec;        // per-image record tables
  unsigned long long* queue;
  unsigned long long qcap;
  uint32_t* flags;    // [nimg]
  uint32_t* done;     // [nimg] completed items
  uint32_t* counter;  // [kCounters]
  double *t, *r, *th, *ph;
  int8_t* rclass;
  int8_t* thclass;
  int32_t* nmax_eq;
  double* tau_tot;
  uint32_t* status;
};

// last image k with key(k) <= key (kind < 0: plan columns, else the first
// item of that pass kind); every lane of a wave searches the same key
__device__ inline int find_image(const BatchImage* imgs, int n, int64_t key, int kind) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t off = kind < 0 ? imgs[mid].slot_off : imgs[mid].item_off[kind];
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
  if (threadIdx.x < kCounters) d.counter[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(kBlock) void kg_batch_setup(BatchDev d) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  // total_slots is a multiple of 64: whole waves leave here together
  if (s >= d.total_slots) return;
  const int im = find_image(d.imgs, d.nimg, s, -1);
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

// Fixup queue item of a batch: walked << 63 | image << 51 | sample << 31 | pixel
// (images < 4096, ngeo <= 2^20, npix < 2^31: checked at plan time).
struct BatchPush {
  unsigned long long* items;
  uint32_t* count;
  unsigned long long cap;
  unsigned long long image;
  __device__ void operator()(int64_t i, int64_t j, bool walked) const {
    const uint32_t s = atomicAdd(count, 1u);
    if (s < cap)
      items[s] = (walked ? (1ull << 63) : 0ull) | (image << 51) | (static_cast<unsigned long long>(j) << 31) |
                 static_cast<unsigned long long>(i);
  }
};

// The radial far pass hands a block with a lost long step to kg_batch_repair
// through the queue storage (the theta fixups are done with it; the radial
// walk's pushes come after the repair pass): at most one item per block and
// pixel slot, within the queue's capacity.
struct BatchRepair {
  static constexpr bool kOn = true;
  BatchPush push;
  __device__ void operator()(int64_t i, int64_t j, bool walked) const { push(i, j, walked); }
};

__device__ inline SampleOut image_out(const BatchDev& d, const BatchImage& B) {
  return SampleOut{nullptr, nullptr, d.t + B.out_off, d.r + B.out_off, d.th + B.out_off, d.ph + B.out_off, nullptr,
                   B.npix, B.ngeo};
}

__device__ inline Records image_records(const BatchDev& d, const BatchImage& B, bool theta) {
  return Records{d.rec + (theta ? B.th_rec : B.r_rec), B.slots, B.slot_off};
}

// One unit of pass KIND (theta if THETA) for pixel i of image B.  Every
// image samples t, phi and sigma (the plan rejects images without
// KGEO_DO_PHI_T), so PHI_T is true throughout.
template <int KIND, bool THETA>
__device__ inline uint32_t batch_unit(const BatchDev& d, const BatchImage& B, int im, int64_t i, int64_t u) {
  const int64_t s = B.slot_off + i;
  const uint32_t fl = d.flags[im];
  const SampleOut o = image_out(d, B);
  const Records rt = image_records(d, B, THETA);
  constexpr bool PHI_T = true;
  if (KIND == K_HEAD) {
    return THETA ? theta_head_pixel<PHI_T>(d.plan, d.total_slots, s, fl, o, i, u, rt)
                 : radial_head_pixel<PHI_T>(B.c, d.plan, d.total_slots, s, fl, o, i, u, rt);
  }
  if (KIND == K_FAR) {
    const BatchRepair repair{
        BatchPush{d.queue, d.counter + kRepairCounter, d.qcap, static_cast<unsigned long long>(im)}};
    return THETA ? theta_far_pixel<PHI_T>(d.plan, d.total_slots, s, fl, o, i, u, rt)
                 : radial_far_pixel<PHI_T>(B.c, d.plan, d.total_slots, s, fl, o, i, u, rt, repair);
  }
  const int64_t j0 = chunk_start(B.ngeo, u, B.units[K_WALK]);
  const int64_t j1 = chunk_start(B.ngeo, u + 1, B.units[K_WALK]);
  BatchPush push{d.queue, d.counter + kQueueCounter + (THETA ? 0 : 1), d.qcap, static_cast<unsigned long long>(im)};
  return THETA ? theta_walk_pixel<PHI_T>(d.plan, d.total_slots, s, fl, o, i, j0, j1, rt, push)
               : radial_walk_pixel<PHI_T>(B.c, d.plan, d.total_slots, s, fl, o, i, j0, j1, rt, push);
}

template <int KIND, bool THETA>
__device__ inline void batch_pass(const BatchDev& d) {
  const int lane = __lane_id();
  uint32_t* counter = d.counter + (THETA ? 0 : K_N) + KIND;
  for (;;) {
    const uint32_t item = fetch_item(counter, lane);
    if (item >= d.total_items[KIND]) break;  // scalar branch: the whole wave leaves once the pass is empty
    const int im = find_image(d.imgs, d.nimg, static_cast<int64_t>(item), KIND);
    const BatchImage& B = d.imgs[im];
    const int64_t li = static_cast<int64_t>(item) - B.item_off[KIND];
    const int64_t units = B.units[KIND];
    const int64_t i = (li / units) * kWave + lane;
    const int64_t u = li % units;
    if (i < B.npix) {
      const uint32_t st = batch_unit<KIND, THETA>(d, B, im, i, u);
      if (d.status && st) atomicOr(&d.status[B.pix_off + i], st);
    }
    atomicAdd(&d.done[im], lane == 0 ? 1u : 0u);
  }
}

template <int KIND, bool THETA>
__global__ __launch_bounds__(kBlock) void kg_batch_pass(BatchDev d) {
  batch_pass<KIND, THETA>(d);
}

// the walk pass held to two waves per SIMD, as kg_radial_walk
template <bool THETA>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void kg_batch_walk(BatchDev d) {
  batch_pass<K_WALK, THETA>(d);
}

// one lane per (queued range, sample of the range) (grid-stride), each
// sample by direct evaluation, as the plain kg_theta_fixup / kg_radial_fixup
template <bool THETA>
__global__ __launch_bounds__(kBlock) void kg_batch_fixup(BatchDev d) {
  const unsigned long long c = d.counter[kQueueCounter + (THETA ? 0 : 1)];
  const unsigned long long n = (c < d.qcap ? c : d.qcap) * kFixupSpan;
  for (unsigned long long e = static_cast<unsigned long long>(blockIdx.x) * kBlock + threadIdx.x; e < n;
       e += static_cast<unsigned long long>(gridDim.x) * kBlock) {
    const unsigned long long q = d.queue[e / kFixupSpan];
    const int64_t sg = static_cast<int64_t>(e % kFixupSpan);
    const int im = static_cast<int>((q >> 51) & 0xfffull);
    const int64_t j = static_cast<int64_t>((q >> 31) & 0xfffffull);
    const int64_t i = static_cast<int64_t>(q & 0x7fffffffull);
    const bool walked = (q >> 63) != 0ull;
    const BatchImage& B = d.imgs[im];
    const int64_t s = B.slot_off + i;
    const uint32_t fl = d.flags[im];
    const SampleOut o = image_out(d, B);
    const uint32_t st = THETA ? theta_fixup_sample<true>(d.plan, d.total_slots, s, fl, o, i, j, walked, sg)
                              : radial_fixup_sample<true>(B.c, d.plan, d.total_slots, s, fl, o, i, j, walked, sg);
    if (d.status && st) atomicOr(&d.status[B.pix_off + i], st);
  }
}

// one lane per (repair item, segment of its block) (grid-stride), as the
// plain kg_radial_repair
constexpr int kRepairSegments = static_cast<int>(kAnchorBlock / kAnchorEvery);

__global__ __launch_bounds__(kBlock) void kg_batch_repair(BatchDev d) {
  const unsigned long long c = d.counter[kRepairCounter];
  const unsigned long long n = (c < d.qcap ? c : d.qcap) * kRepairSegments;
  for (unsigned long long e = static_cast<unsigned long long>(blockIdx.x) * kBlock + threadIdx.x; e < n;
       e += static_cast<unsigned long long>(gridDim.x) * kBlock) {
    const unsigned long long q = d.queue[e / kRepairSegments];
    const int64_t sg = static_cast<int64_t>(e % kRepairSegments);
    const int im = static_cast<int>((q >> 51) & 0xfffull);
    const int64_t j = static_cast<int64_t>((q >> 31) & 0xfffffull);
    const int64_t i = static_cast<int64_t>(q & 0x7fffffffull);
    const BatchImage& B = d.imgs[im];
    const int64_t s = B.slot_off + i;
    const uint32_t fl = d.flags[im];
    const SampleOut o = image_out(d, B);
    const Records rt = image_records(d, B, false);
    const uint32_t st = radial_repair_pixel<true>(B.c, d.plan, d.total_slots, s, fl, o, i, j, sg, rt);
    if (d.status && st) atomicOr(&d.status[B.pix_off + i], st);
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
  if (nimg > 4096) return bfail(KGEO_E_ARG, "more than 4096 images (%lld)", nimg);
  for (int32_t k = 0; k < nimg; ++k) {
    if (imgs[k].npix < 1) return bfail(KGEO_E_ARG, "image %lld has npix < 1", k);
    if (imgs[k].npix >= (1LL << 31)) return bfail(KGEO_E_ARG, "image %lld has npix >= 2^31", k);
    if (imgs[k].ngeo < 1) return bfail(KGEO_E_ARG, "image %lld has ngeo < 1", k);
    if (imgs[k].ngeo > (1LL << 20)) return bfail(KGEO_E_ARG, "image %lld has ngeo > 2^20", k);
    // every image samples t, phi and sigma, as the plain kernels do
    if (!(imgs[k].flags & KGEO_DO_PHI_T)) return bfail(KGEO_E_ARG, "image %lld lacks KGEO_DO_PHI_T", k);
  }
  return KGEO_OK;
}

int64_t slots_of(const kgeo_ray_params& p) { return (p.npix + kWave - 1) / kWave * kWave; }

int64_t total_slots_of(const kgeo_ray_params* imgs, int32_t nimg) {
  int64_t s = 0;
  for (int32_t k = 0; k < nimg; ++k) s += slots_of(imgs[k]);
  return s;
}

// Elements of image k's theta and radial record tables (a walked grid needs
// ngeo > 2; smaller images keep none) and its fixup-queue share.
int64_t theta_rec_elems(const kgeo_ray_params& p) {
  return p.ngeo > 2 ? theta_records(p.ngeo) * kThFields * slots_of(p) : 0;
}
int64_t radial_rec_elems(const kgeo_ray_params& p) {
  return p.ngeo > 2 ? radial_records(p.ngeo) * kRFields * slots_of(p) : 0;
}
int64_t queue_elems(const kgeo_ray_params& p) { return fixup_items(p.ngeo) * slots_of(p); }

struct BatchSizes {
  int64_t slots, rec, queue;
};

BatchSizes sizes_of(const kgeo_ray_params* imgs, int32_t nimg) {
  BatchSizes z{0, 0, 0};
  for (int32_t k = 0; k < nimg; ++k) {
    z.slots += slots_of(imgs[k]);
    z.rec += theta_rec_elems(imgs[k]) + radial_rec_elems(imgs[k]);
    z.queue += queue_elems(imgs[k]);
  }
  return z;
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
  const BatchSizes z = sizes_of(imgs, nimg);
  return align256(sizeof(BatchImage) * static_cast<size_t>(nimg)) +
         2 * align256(sizeof(uint32_t) * static_cast<size_t>(nimg)) + 256 +
         align256(static_cast<size_t>(z.slots) * PL_N * sizeof(double)) +
         align256(static_cast<size_t>(z.rec) * sizeof(double)) +
         static_cast<size_t>(z.queue) * sizeof(unsigned long long);
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
  int64_t pix = 0, slots = 0, outo = 0, rec = 0;
  int64_t items[K_N] = {0, 0, 0};
  // walk items hold whole segments (chunk starts are multiples of kAnchorEveryLong)
  const int64_t spi =
      (static_cast<int64_t>(tn.samples_per_item) + kAnchorEveryLong - 1) / kAnchorEveryLong * kAnchorEveryLong;
  for (int32_t k = 0; k < nimg; ++k) {
    const kgeo_ray_params& p = imgs[k];
    BatchImage& B = host[k];
    B.c = scalars_of(p);
    B.npix = p.npix;
    B.ngeo = p.ngeo;
    B.pix_off = pix;
    B.slot_off = slots;
    B.slots = slots_of(p);
    B.out_off = outo;
    B.units[K_HEAD] = static_cast<int32_t>(anchor_heads(p.ngeo));
    B.units[K_FAR] = static_cast<int32_t>(anchor_blocks(p.ngeo));
    B.units[K_WALK] = static_cast<int32_t>((p.ngeo + spi - 1) / spi);
    B.th_rec = rec;
    rec += theta_rec_elems(p);
    B.r_rec = rec;
    rec += radial_rec_elems(p);
    B.do_phi_t = 1;  // required by check_images
    const int64_t tiles = B.slots / kWave;
    int64_t mine = 0;
    for (int q = 0; q < K_N; ++q) {
      B.item_off[q] = items[q];
      items[q] += tiles * B.units[q];
      mine += tiles * B.units[q];
      if (items[q] > kMaxItems) return bfail(KGEO_E_ARG, "too many work items (%lld); raise samples_per_item", items[q]);
    }
    pix += p.npix;
    slots += B.slots;
    outo += p.npix * p.ngeo;
    per[k] = static_cast<uint32_t>(2 * mine);  // theta and radial passes
  }
  if (!t || !ph) return bfail(KGEO_E_ARG, "t and ph are required");

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
  w += align256(static_cast<size_t>(slots) * PL_N * sizeof(double));
  d.rec = reinterpret_cast<double*>(w);
  w += align256(static_cast<size_t>(rec) * sizeof(double));
  d.queue = reinterpret_cast<unsigned long long*>(w);
  d.qcap = static_cast<unsigned long long>(sizes_of(imgs, nimg).queue);
  d.nimg = nimg;
  d.total_slots = slots;
  for (int q = 0; q < K_N; ++q) d.total_items[q] = static_cast<uint32_t>(items[q]);
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
  hipLaunchKernelGGL((kg_batch_pass<K_HEAD, true>), dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL((kg_batch_pass<K_FAR, true>), dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(kg_batch_walk<true>, dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(kg_batch_fixup<true>, dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL((kg_batch_pass<K_HEAD, false>), dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL((kg_batch_pass<K_FAR, false>), dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(kg_batch_repair, dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(kg_batch_walk<false>, dim3(blocks), dim3(kBlock), 0, st, d);
  hipLaunchKernelGGL(kg_batch_fixup<false>, dim3(blocks), dim3(kBlock), 0, st, d);
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

int64_t kgeo_batch_items(const kgeo_batch* b) {
  if (!b) return -1;
  int64_t n = 0;
  for (int q = 0; q < K_N; ++q) n += 2 * static_cast<int64_t>(b->d.total_items[q]);
  return n;
}

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
