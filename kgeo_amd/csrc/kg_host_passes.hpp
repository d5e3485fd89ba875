// The head / far / walk / fixup passes of the sample kernels (kg_sample.hpp,
// kgeo_raytrace.hip) run on the CPU over all pixels, in the kernels' order:
// theta heads, long steps, walk, fixups, then the same for r, whose long
// steps are followed by the repair of blocks with a lost long step.  Used by the
// CPU test build (T = double) and the FLOP counter (T = its counting type).

#pragma once

#include <cstdint>
#include <vector>

#include "kg_sample.hpp"

namespace kg {

// Fixup queue on the host: the kernels' 64-bit item (walked << 63 |
// first sample << 32 | pixel).
struct HostPush {
  std::vector<uint64_t>* items;
  void operator()(int64_t i, int64_t j, bool walked) const {
    items->push_back((walked ? (1ull << 63) : 0ull) | (static_cast<uint64_t>(j) << 32) | static_cast<uint32_t>(i));
  }
};

// The far pass's repair items, on the host (the same item format).
struct HostRepair {
  static constexpr bool kOn = true;
  std::vector<uint64_t>* items;
  void operator()(int64_t i, int64_t j, bool walked) const { HostPush{items}(i, j, walked); }
};

inline void unpack_host_item(uint64_t e, int64_t& i, int64_t& j, bool& walked) {
  i = static_cast<int64_t>(e & 0xffffffffull);
  j = static_cast<int64_t>((e >> 32) & 0x7fffffffull);
  walked = (e >> 63) != 0ull;
}

// st [npix] accumulates the status bits; nfix[3] receives the theta and
// radial queue lengths, or -1 where a pass queued more items than the
// kernels' queue holds (npix * fixup_items(ngeo)), and the number of
// radial repair items.  The host queue
// itself is unbounded, so the results stay complete either way.
template <bool DO_PHI_T, class T>
void three_passes(const RayScalars<T>& c, const double* plan, uint32_t gflags, const SampleOut& o, int chunks,
                  uint32_t* st, int64_t* nfix) {
  const int64_t npix = o.npix, G = o.ngeo;
  std::vector<double> trec(static_cast<size_t>(theta_records(G)) * kThFields * npix);
  std::vector<double> rrec(static_cast<size_t>(radial_records(G)) * kRFields * npix);
  const Records rth{trec.data(), npix, 0}, rr{rrec.data(), npix, 0};
  std::vector<uint64_t> items;
  HostPush push{&items};
  const int64_t nblk = anchor_blocks(G), nhead = anchor_heads(G);
  for (int64_t i = 0; i < npix; ++i)
    for (int64_t hd = 0; hd < nhead; ++hd) st[i] |= theta_head_pixel<DO_PHI_T, T>(plan, npix, i, gflags, o, i, hd, rth);
  for (int64_t i = 0; i < npix; ++i)
    for (int64_t b = 0; b < nblk; ++b) st[i] |= theta_far_pixel<DO_PHI_T, T>(plan, npix, i, gflags, o, i, b, rth);
  for (int64_t i = 0; i < npix; ++i)
    for (int ch = 0; ch < chunks; ++ch)
      st[i] |= theta_walk_pixel<DO_PHI_T, HostPush, T>(plan, npix, i, gflags, o, i, chunk_start(G, ch, chunks),
                                                        chunk_start(G, ch + 1, chunks), rth, push);
  const size_t cap = static_cast<size_t>(npix) * static_cast<size_t>(fixup_items(G));
  nfix[0] = items.size() <= cap ? static_cast<int64_t>(items.size()) : -1;
  for (uint64_t e : items) {
    int64_t i, j;
    bool w;
    unpack_host_item(e, i, j, w);
    for (int64_t sg = 0; sg < kFixupSpan; ++sg)
      st[i] |= theta_fixup_sample<DO_PHI_T, T>(plan, npix, i, gflags, o, i, j, w, sg);
  }
  items.clear();
  for (int64_t i = 0; i < npix; ++i)
    for (int64_t hd = 0; hd < nhead; ++hd)
      st[i] |= radial_head_pixel<DO_PHI_T, T>(c, plan, npix, i, gflags, o, i, hd, rr);
  std::vector<uint64_t> repairs;
  const HostRepair repair{&repairs};
  for (int64_t i = 0; i < npix; ++i)
    for (int64_t b = 0; b < nblk; ++b)
      st[i] |= radial_far_pixel<DO_PHI_T, T>(c, plan, npix, i, gflags, o, i, b, rr, repair);
  nfix[2] = static_cast<int64_t>(repairs.size());
  for (uint64_t e : repairs) {
    int64_t i, j;
    bool w;
    unpack_host_item(e, i, j, w);
    for (int64_t sg = 0; sg < kAnchorBlock / kAnchorEvery; ++sg)
      st[i] |= radial_repair_pixel<DO_PHI_T, T>(c, plan, npix, i, gflags, o, i, j, sg, rr);
  }
  for (int64_t i = 0; i < npix; ++i)
    for (int ch = 0; ch < chunks; ++ch)
      st[i] |= radial_walk_pixel<DO_PHI_T, HostPush, T>(c, plan, npix, i, gflags, o, i, chunk_start(G, ch, chunks),
                                                         chunk_start(G, ch + 1, chunks), rr, push);
  nfix[1] = items.size() <= cap ? static_cast<int64_t>(items.size()) : -1;
  for (uint64_t e : items) {
    int64_t i, j;
    bool w;
    unpack_host_item(e, i, j, w);
    for (int64_t sg = 0; sg < kFixupSpan; ++sg)
      st[i] |= radial_fixup_sample<DO_PHI_T, T>(c, plan, npix, i, gflags, o, i, j, w, sg);
  }
}

}  // namespace kg
