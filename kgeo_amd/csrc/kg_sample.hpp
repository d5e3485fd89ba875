// Per-pixel sample loops shared by the plain kernels (kgeo_raytrace.hip) and
// the persistent batch kernel (kgeo_batch.hip), so both run the same code.
//
// theta_pixel: loads the theta slots of one pixel's plan, builds the theta
//   Jacobi ladder and, for samples j0..j1-1, writes theta and parks
//   G_phi -> ph, G_t -> t (kerr_raytracing_ana.py:186-188 needs them there).
// radial_pixel: loads the radial slots, builds the radial ladder, writes r and
//   combines phi = I_phi + lambda G_phi, t = I_t + a^2 G_t,
//   sigma = I_sigma + a^2 G_t (kerr_raytracing_ana.py:186-188).
// Each runs one loop specialised at compile time for the pixel's motion class
// or radial case, so a wave only keeps its own case's constants live.  The
// plan is slot-major: slot s of pixel p is plan[s * stride + p].
//
// On the linspace grid the loops walk: each sample's state follows from the
// previous one by the addition theorems (kg_step.hpp), with a direct
// evaluation at the chunk's first sample, at every multiple of the anchor
// interval (kg_step.hpp),
// at the last grid sample (tau = stop exactly) and wherever a step is not
// well conditioned.  With caller-given Mino times (tau_in) every sample is direct.
// Host-compilable (KG_INLINE), so the CPU test build runs the same loops.

#pragma once

#include "kg_geodesic.hpp"
#include "kg_step.hpp"

namespace kg {

enum : uint32_t { GF_ANY_ETA_POS = 1u, GF_ANY_STEP_ZERO = 2u };

struct SampleOut {
  const double* tau_in;  // optional [ngeo][npix] Mino times replacing the grid
  double* tau_out;       // optional
  double* t;
  double* r;
  double* th;
  double* ph;
  double* sig;           // optional
  int64_t npix;          // row stride of the [ngeo][npix] arrays
  int64_t ngeo;
};

KG_INLINE double sample_tau(const SampleOut& o, int64_t j, int64_t idx, double stop, bool any_step_zero) {
  return o.tau_in ? o.tau_in[idx] : mino_tau(j, o.ngeo, stop, any_step_zero);
}

// The loops below are templated on the scalar type: T = double in the
// kernels; the op-counting CPU build instantiates them on its counting type
// (outputs are stored through val()).

// Writers shared by the direct and the walking loops.
template <class T, bool DO_PHI_T>
KG_INLINE bool theta_store(const SampleOut& o, int64_t idx, double tau, const T& th, const T& Gph, const T& Gt) {
  if (o.tau_out) o.tau_out[idx] = tau;
  if (o.th) o.th[idx] = val(th);
  if (DO_PHI_T) {
    o.ph[idx] = val(Gph);
    o.t[idx] = val(Gt);
  }
  return !kfinite(val(th));
}

// Gph, Gt: the theta pass's G_phi, G_t of this sample (parked in ph, t), when
// the caller has loaded them ahead of the arithmetic (radial_store_g).
template <class T, bool DO_PHI_T>
KG_INLINE bool radial_store_g(const SampleOut& o, const CombineConsts<T>& cmb, int64_t idx, double tau,
                              const RadialOut<T>& ro, const T& Gph, const T& Gt) {
  if (o.r) o.r[idx] = val(ro.r);
  bool bad = !kfinite(val(ro.r));
  if (DO_PHI_T) {
    Sample<T> s;
    combine(cmb, T(tau), ro, T(0.0), Gph, Gt, s);
    o.t[idx] = val(s.t);
    o.ph[idx] = val(s.ph);
    if (o.sig) o.sig[idx] = val(s.sig);
    bad |= !(kfinite(val(s.t)) && kfinite(val(s.ph)) && kfinite(val(s.sig)));
  }
  return bad;
}

template <class T, bool DO_PHI_T>
KG_INLINE bool radial_store(const SampleOut& o, const CombineConsts<T>& cmb, int64_t idx, double tau,
                            const RadialOut<T>& ro) {
  if (o.r) o.r[idx] = val(ro.r);
  bool bad = !kfinite(val(ro.r));
  if (DO_PHI_T) {
    const T Gph = T(o.ph[idx]);
    const T Gt = T(o.t[idx]);
    Sample<T> s;
    combine(cmb, T(tau), ro, T(0.0), Gph, Gt, s);
    o.t[idx] = val(s.t);
    o.ph[idx] = val(s.ph);
    if (o.sig) o.sig[idx] = val(s.sig);
    bad |= !(kfinite(val(s.t)) && kfinite(val(s.ph)) && kfinite(val(s.sig)));
  }
  return bad;
}

// Direct evaluation of every sample (caller-given Mino times, numpy's
// zero-step grid, and the motion classes and cases the walk does not cover).
template <class T, int KIND, bool DO_PHI_T>
KG_INLINE bool theta_loop(const SampleOut& o, const ThetaConsts<T>& tc, const JacobiLadder<T>& lad, int64_t i,
                          int64_t j0, int64_t j1, double stop, bool any_step_zero) {
  bool bad = false;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t idx = j * o.npix + i;
    const double tau = sample_tau(o, j, idx, stop, any_step_zero);
    T th = T(0.0), Gph = T(0.0), Gt = T(0.0);
    if (KIND != THK_NONE) theta_eval<T, KIND>(tc, lad, T(tau), DO_PHI_T, th, Gph, Gt);
    bad |= theta_store<T, DO_PHI_T>(o, idx, tau, th, Gph, Gt);
  }
  return bad;
}

template <class T, int CASE, bool DO_PHI_T>
KG_INLINE bool radial_loop(const SampleOut& o, const RadialConsts<T>& rc, const JacobiLadder<T>& lad,
                           const CombineConsts<T>& cmb, int64_t i, int64_t j0, int64_t j1, double stop,
                           bool any_step_zero) {
  bool bad = false;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t idx = j * o.npix + i;
    const double tau = sample_tau(o, j, idx, stop, any_step_zero);
    RadialOut<T> ro = radial_eval<T, CASE>(rc, lad, T(tau), DO_PHI_T);
    bad |= radial_store<T, DO_PHI_T>(o, cmb, idx, tau, ro);
  }
  return bad;
}

// Theta samples j0..j1-1 on the linspace grid (ordinary or vortical motion),
// walking u_j = w (s_G - j h) with the addition theorems.
template <class T, int KIND, bool DO_PHI_T>
KG_INLINE bool theta_loop_walk(const SampleOut& o, const ThetaConsts<T>& tc, const JacobiLadder<T>& lad, int64_t i,
                               int64_t j0, int64_t j1, double stop) {
  const T m = tc.v[TH_k];
  const double h = stop / static_cast<double>(o.ngeo - 1);
  WalkStep<T, 1> w;
  {
    ThState<T> sb;
    const T du = -tc.v[TH_w] * h;
    theta_state<T, KIND>(tc, lad, du, DO_PHI_T, sb);
    const T n[1] = {tc.v[TH_n]};
    walk_init(w, m, n);
    w.b.s = sb.s;
    w.b.c = sb.c;
    w.b.d = sb.d;
    w.b.D = DO_PHI_T ? sb.EmF / m : T(0.0);
    if (DO_PHI_T) walk_enter(w, 0, sb.Pi, du, sb.s, sb.c, sb.d, w.b);
  }
  WalkState<T, 1> st;
  bool bad = false;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t idx = j * o.npix + i;
    const double tau = mino_tau(j, o.ngeo, stop, false);
    const T snarg = theta_arg(tc, T(tau));
    ThState<T> ts;
    bool direct = (j == j0) || (j % kAnchorEveryLong == 0) || (j == o.ngeo - 1);
    if (!direct) {
      WalkState<T, 1> nx;
      if (walk_step<T, 1, DO_PHI_T>(st, w, nx)) {
        st = nx;
        ts.s = st.s;
        ts.c = st.c;
        ts.d = st.d;
        ts.F = snarg;
        ts.EmF = m * st.D;
        ts.Pi = DO_PHI_T ? walk_pi(w, 0, st, snarg) : T(0.0);
      } else {
        direct = true;
      }
    }
    if (direct) {
      theta_state<T, KIND>(tc, lad, snarg, DO_PHI_T, ts);
      st.s = ts.s;
      st.c = ts.c;
      st.d = ts.d;
      st.D = DO_PHI_T ? ts.EmF / m : T(0.0);
      if (DO_PHI_T) walk_enter(w, 0, ts.Pi, snarg, ts.s, ts.c, ts.d, st);
    }
    T th = T(0.0), Gph = T(0.0), Gt = T(0.0);
    theta_out<T, KIND>(tc, snarg, ts, DO_PHI_T, th, Gph, Gt);
    bad |= theta_store<T, DO_PHI_T>(o, idx, tau, th, Gph, Gt);
  }
  return bad;
}

// A direct block log entering the walk: next to its singularity its rounding
// error (~eps e^|lg|) would be carried along the segment, so it enters as
// NaN and the next step is evaluated directly.
template <class T> KG_INLINE T block_log_enter(const T& lg) {
  return kfabs(val(lg)) <= kLogGuard ? lg : T(NAN);
}

// One walked step of the case 3 / 4 block logs from state a to state b;
// false (lb then unspecified) where a log leaves [-kLogGuard, kLogGuard].
// ADAPT as in walk_step.
template <class T, int CASE, bool ADAPT = false>
KG_INLINE bool radial_block_step(const RadialConsts<T>& rc, const WalkState<T, 3>& a, const WalkState<T, 3>& b,
                                 T (&lb)[3]) {
  T a0[3], b0[3], a1[3], b1[3], den[3], iden[3], y[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    block_log_args<T, CASE>(rc, q, a.s, a.d, a0[q], b0[q]);
    block_log_args<T, CASE>(rc, q, b.s, b.d, a1[q], b1[q]);
    den[q] = walk_log_den(a0[q], b0[q], a1[q], b1[q]);
  }
  shared_reciprocals<T, 3>(den, iden);
  double ymax = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    y[q] = walk_log_y(a0[q], b0[q], a1[q], b1[q], iden[q]);
    const double y2 = kfabs(val(y[q] * y[q]));
    ymax = kmax(ymax, y2 <= kStepZ2 ? y2 : 0.0);
  }
  const int nt = rc_terms<ADAPT>(ymax);
  bool ok = true;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    lb[q] = walk_log(lb[q], y[q], a1[q], b1[q], nt);
    ok = ok & (kfabs(val(lb[q])) <= kLogGuard);
  }
  return ok;
}

// Radial samples j0..j1-1 on the linspace grid, walking X_j = U0 - Xs j h.
template <class T, int CASE, bool DO_PHI_T>
KG_INLINE bool radial_loop_walk(const SampleOut& o, const RadialConsts<T>& rc, const JacobiLadder<T>& lad,
                                const CombineConsts<T>& cmb, int64_t i, int64_t j0, int64_t j1, double stop) {
  const T m = rc.v[R_k];
  const double h = stop / static_cast<double>(o.ngeo - 1);
  WalkStep<T, 3> w;
  {
    const T dX = -rc.v[R_Xs] * h;
    RState<T> sb;
    radial_state<T, CASE>(rc, lad, dX, T(INFINITY), DO_PHI_T, sb);  // plain Jacobi at dX
    T n[3], pc[3];
    radial_pi_params<T, CASE>(rc, n, pc);
    walk_init(w, m, n);
    w.b.s = sb.s;
    w.b.c = sb.c;
    w.b.d = sb.d;
    w.b.D = DO_PHI_T ? sb.EmF / m : T(0.0);
#pragma unroll
    for (int q = 0; q < 3; ++q)
      if (DO_PHI_T) walk_enter(w, q, sb.Pi[q], dX, sb.s, sb.c, sb.d, w.b);
  }
  constexpr bool kBlocks = DO_PHI_T && (CASE == 3 || CASE == 4);
  // case 3 and 4 outputs amplify the walk's rounding (1/(1 - al^2) in the R1
  // blocks), so they restart from a direct evaluation twice as often
  constexpr int kAnchor = (CASE == 2) ? kAnchorEveryLong : kAnchorEvery;
  WalkState<T, 3> st;
  T Lb[3];  // cases 3, 4: walked logs of the R1 / S1 blocks
  bool bad = false;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t idx = j * o.npix + i;
    const double tau = mino_tau(j, o.ngeo, stop, false);
    const T dv = rc.v[R_Xs] * tau;
    const T X = rc.v[R_U0] - dv;
    RState<T> rs;
    bool direct = (j == j0) || (j % kAnchor == 0) || (j == o.ngeo - 1);
    if (!direct) {
      WalkState<T, 3> nx;
      T lb1[3] = {Lb[0], Lb[1], Lb[2]};
      if (walk_step<T, 3, DO_PHI_T>(st, w, nx) && (!kBlocks || radial_block_step<T, CASE>(rc, st, nx, lb1))) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          Lb[q] = lb1[q];
          rs.lg[q] = lb1[q];
        }
        st = nx;
        rs.s = st.s;
        rs.c = st.c;
        rs.d = st.d;
        rs.F = X;
        rs.EmF = m * st.D;
#pragma unroll
        for (int q = 0; q < 3; ++q) rs.Pi[q] = DO_PHI_T ? walk_pi(w, q, st, X) : T(0.0);
      } else {
        direct = true;
      }
    }
    if (direct) {
      radial_state<T, CASE>(rc, lad, X, dv, DO_PHI_T, rs);
      st.s = rs.s;
      st.c = rs.c;
      st.d = rs.d;
      st.D = DO_PHI_T ? rs.EmF / m : T(0.0);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (DO_PHI_T) walk_enter(w, q, rs.Pi[q], X, rs.s, rs.c, rs.d, st);
        if (kBlocks) Lb[q] = block_log_enter(rs.lg[q]);
      }
    }
    RadialOut<T> ro = radial_out<T, CASE>(rc, T(tau), rs, DO_PHI_T);
    bad |= radial_store<T, DO_PHI_T>(o, cmb, idx, tau, ro);
  }
  return bad;
}

// First sample of chunk c of `chunks` over ngeo samples.  Chunk starts are
// multiples of kAnchorEveryLong (itself a multiple of kAnchorEvery), so the
// walk's direct evaluations fall on the same samples (the anchor multiples,
// the last sample and the ill-conditioned steps) for every launch shape:
// results are bit-identical whatever the chunking.
KG_INLINE int64_t chunk_start(int64_t ngeo, int64_t c, int64_t chunks) {
  if (c <= 0) return 0;
  if (c >= chunks) return ngeo;
  return ((ngeo * c) / chunks) / kAnchorEveryLong * kAnchorEveryLong;
}

// Walking needs the linspace grid with at least one interior step: not for
// caller-given Mino times and not for numpy's zero-step branch.
KG_INLINE bool walk_ok(const SampleOut& o, bool any_step_zero) {
  return o.tau_in == nullptr && !any_step_zero && o.ngeo > 2;
}

template <class T> KG_INLINE uint32_t loop_status(bool bad, const JacobiLadder<T>& lad) {
  return (bad ? ST_NONFINITE : 0u) | (lad.capped ? ST_LADDER_CAP : 0u);
}

// Theta part of samples j0..j1-1 of output pixel i, whose plan column is p.
// WALK = false: every sample direct (the kernels for caller-given Mino times
// and grids of at most two samples, which cannot walk).
template <bool DO_PHI_T, class T = double, bool WALK = true>
KG_INLINE uint32_t theta_pixel(const double* __restrict__ plan, int64_t stride, int64_t p, uint32_t gflags,
                               const SampleOut& o, int64_t i, int64_t j0, int64_t j1) {
  ThetaConsts<T> tc;
#pragma unroll
  for (int s = 0; s < TH_NSLOT; ++s) tc.v[s] = T(plan[static_cast<int64_t>(PL_TH + s) * stride + p]);
  const int kind = static_cast<int>(plan[static_cast<int64_t>(PL_THK) * stride + p]);
  const bool eta0 = plan[static_cast<int64_t>(PL_ETA0) * stride + p] != 0.0;
  const double stop = plan[static_cast<int64_t>(PL_STOP) * stride + p];
  const bool any_eta_pos = (gflags & GF_ANY_ETA_POS) != 0;
  const bool any_step_zero = (gflags & GF_ANY_STEP_ZERO) != 0;
  // eta == 0 pixels take theta = 0 when no pixel of the image has eta > 0
  const int k = (eta0 && !any_eta_pos) ? THK_NONE : kind;
  JacobiLadder<T> lad;
  jacobi_ladder(tc.v[TH_mp], lad);
  const bool walk = WALK && walk_ok(o, any_step_zero);
  bool bad;
  switch (k) {
    case THK_ORD:
      bad = walk ? theta_loop_walk<T, THK_ORD, DO_PHI_T>(o, tc, lad, i, j0, j1, stop)
                 : theta_loop<T, THK_ORD, DO_PHI_T>(o, tc, lad, i, j0, j1, stop, any_step_zero);
      break;
    case THK_VORT:
      bad = walk ? theta_loop_walk<T, THK_VORT, DO_PHI_T>(o, tc, lad, i, j0, j1, stop)
                 : theta_loop<T, THK_VORT, DO_PHI_T>(o, tc, lad, i, j0, j1, stop, any_step_zero);
      break;
    case THK_ORD_K0: bad = theta_loop<T, THK_ORD_K0, DO_PHI_T>(o, tc, lad, i, j0, j1, stop, any_step_zero); break;
    default: bad = theta_loop<T, THK_NONE, DO_PHI_T>(o, tc, lad, i, j0, j1, stop, any_step_zero); break;
  }
  return loop_status(bad, lad);
}

// Radial part (and the phi, t, sigma combination) of the same samples.
template <bool DO_PHI_T, class T = double, bool WALK = true>
KG_INLINE uint32_t radial_pixel(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride, int64_t p,
                                uint32_t gflags, const SampleOut& o, int64_t i, int64_t j0, int64_t j1) {
  RadialConsts<T> rc;
#pragma unroll
  for (int s = 0; s < R_NSLOT; ++s) rc.v[s] = T(plan[static_cast<int64_t>(PL_R + s) * stride + p]);
  rc.rcase = static_cast<int>(plan[static_cast<int64_t>(PL_RCASE) * stride + p]);
  const T lam = T(plan[static_cast<int64_t>(PL_LAM) * stride + p]);
  const double stop = plan[static_cast<int64_t>(PL_STOP) * stride + p];
  const bool any_step_zero = (gflags & GF_ANY_STEP_ZERO) != 0;
  const CombineConsts<T> cmb = combine_consts(c, lam);
  JacobiLadder<T> lad;
  jacobi_ladder(rc.v[R_k], lad);
  const bool walk = WALK && walk_ok(o, any_step_zero) && lad.n >= 0;
  bool bad;
  switch (rc.rcase) {
    case 4:
      bad = walk ? radial_loop_walk<T, 4, DO_PHI_T>(o, rc, lad, cmb, i, j0, j1, stop)
                 : radial_loop<T, 4, DO_PHI_T>(o, rc, lad, cmb, i, j0, j1, stop, any_step_zero);
      break;
    case 3:
      bad = walk ? radial_loop_walk<T, 3, DO_PHI_T>(o, rc, lad, cmb, i, j0, j1, stop)
                 : radial_loop<T, 3, DO_PHI_T>(o, rc, lad, cmb, i, j0, j1, stop, any_step_zero);
      break;
    default:
      bad = walk ? radial_loop_walk<T, 2, DO_PHI_T>(o, rc, lad, cmb, i, j0, j1, stop)
                 : radial_loop<T, 2, DO_PHI_T>(o, rc, lad, cmb, i, j0, j1, stop, any_step_zero);
      break;
  }
  return loop_status(bad, lad);
}

// ---------------------------------------------------------------------------
// The walk in passes, each its own kernel, so the walking loop is not held
// to the register budget of the direct evaluation (Landen ladder, Carlson
// duplication for three characteristics), which it needs only where it
// restarts:
//   heads   one lane per (pixel, head): the direct evaluation at the grid
//           step h and at the segment length L h (records 0 and 1), at the
//           last grid sample and at the first sample of each block of
//           kAnchorBlock samples; stores the walk state in the record table
//           and, for a sample, its outputs;
//   far     one lane per (pixel, block): the block's later segment starts,
//           each one addition step of L samples from the previous (record 1,
//           R_C in closed form where the series does not reach); a refused
//           step marks the block's remaining records unreachable;
//   walk    one lane per (pixel, sample chunk): every other sample from the
//           addition theorems alone, restarting from the records; a refused
//           step (or an unreachable record) queues the rest of its segment
//           (in ranges of at most kFixupSpan samples) for the fixup pass and
//           the lane idles to the next segment;
//   fixup   one lane per queued (pixel, range): the combined loops above
//           (direct evaluation where the walk cannot step) over the range.  Pixels the walk does not cover
//           (theta classes other than ordinary and vortical, radial m >= 1,
//           numpy's zero-step grid, ngeo <= 2) queue all their segments.
// A segment is kAnchorEveryLong samples (theta, radial case 2) or kAnchorEvery
// samples (radial cases 3 and 4) starting at a multiple of its length; blocks
// start at multiples of kAnchorBlock.  Both are fixed by the sample index, so
// every sample is computed the same way whatever the chunking.  A sample is
// at most kAnchorBlock / L - 1 long steps and L - 1 grid steps from a direct
// evaluation.  Records are slot-major [record][field][stride]: 0 the step h,
// 1 the step L h, 2 + k the sample k * (theta: kAnchorEveryLong, radial:
// kAnchorEvery), the last record the sample ngeo - 1.
constexpr int kThFields = 6;   // s, c, d, D, P, Lg
constexpr int kRFields = 13;   // s, c, d, D, P[3], Lg[3], Lb[3] (walked R1/S1 logs)
constexpr int64_t kAnchorBlock = 256;

KG_INLINE int64_t seg_count(int64_t ngeo, int64_t len) { return (ngeo + len - 1) / len; }

// A queued range ends at the next multiple of kFixupSpan (or the segment's
// end): the fixup pass runs its ranges in parallel, and a range of refused
// steps costs a direct evaluation per sample.  Each kFixupSpan window of a
// pixel is queued at most once, so ceil(ngeo / kFixupSpan) items per pixel
// bound the queue.
constexpr int64_t kFixupSpan = 8;

KG_INLINE int64_t fixup_items(int64_t ngeo) { return (ngeo + kFixupSpan - 1) / kFixupSpan; }

// Queue walked samples j .. min(segend, last) - 1 (the last grid sample is
// the anchor pass's): from j to the next window start, then one item per
// window.
template <class Push>
KG_INLINE void push_range(Push& push, int64_t i, int64_t j, int64_t segend, int64_t last) {
  const int64_t e = segend < last ? segend : last;
  push(i, j, true);
  for (int64_t k = (j / kFixupSpan + 1) * kFixupSpan; k < e; k += kFixupSpan) push(i, k, true);
}
KG_INLINE int64_t theta_records(int64_t ngeo) { return seg_count(ngeo, kAnchorEveryLong) + 3; }
KG_INLINE int64_t radial_records(int64_t ngeo) { return seg_count(ngeo, kAnchorEvery) + 3; }
KG_INLINE int64_t anchor_blocks(int64_t ngeo) { return seg_count(ngeo, kAnchorBlock); }

// Column col of the table is the plan column col0 + col - col0: a table of
// one image of a batch starts at that image's first plan column col0.
struct Records {
  double* p;
  int64_t stride;
  int64_t col0;
  KG_INLINE double* at(int64_t rec, int nf, int f, int64_t col) const {
    return p + (rec * nf + f) * stride + (col - col0);
  }
};

template <class T, int NP>
KG_INLINE void record_store(const Records& rt, int64_t rec, int nf, int64_t col, const WalkState<T, NP>& st,
                            const T* lb) {
  *rt.at(rec, nf, 0, col) = val(st.s);
  *rt.at(rec, nf, 1, col) = val(st.c);
  *rt.at(rec, nf, 2, col) = val(st.d);
  *rt.at(rec, nf, 3, col) = val(st.D);
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    *rt.at(rec, nf, 4 + q, col) = val(st.P[q]);
    *rt.at(rec, nf, 4 + NP + q, col) = val(st.Lg[q]);
    if (lb) *rt.at(rec, nf, 4 + 2 * NP + q, col) = val(lb[q]);
  }
}

template <class T, int NP>
KG_INLINE void record_load(const Records& rt, int64_t rec, int nf, int64_t col, WalkState<T, NP>& st, T* lb) {
  st.s = T(*rt.at(rec, nf, 0, col));
  st.c = T(*rt.at(rec, nf, 1, col));
  st.d = T(*rt.at(rec, nf, 2, col));
  st.D = T(*rt.at(rec, nf, 3, col));
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    st.P[q] = T(*rt.at(rec, nf, 4 + q, col));
    st.Lg[q] = T(*rt.at(rec, nf, 4 + NP + q, col));
    if (lb) lb[q] = T(*rt.at(rec, nf, 4 + 2 * NP + q, col));
  }
}

// The last grid sample's radial record: r, I_+, I_-, I_1, I_2 of its direct
// evaluation (fields 0..4), which the walk combines with the theta pass's
// G_phi, G_t.
template <class T>
KG_INLINE void last_out_store(const Records& rt, int64_t rec, int64_t col, const RadialOut<T>& ro) {
  *rt.at(rec, kRFields, 0, col) = val(ro.r);
  *rt.at(rec, kRFields, 1, col) = val(ro.Ip);
  *rt.at(rec, kRFields, 2, col) = val(ro.Im);
  *rt.at(rec, kRFields, 3, col) = val(ro.I1);
  *rt.at(rec, kRFields, 4, col) = val(ro.I2);
}

template <class T>
KG_INLINE RadialOut<T> last_out_load(const Records& rt, int64_t rec, int64_t col) {
  RadialOut<T> ro;
  ro.r = T(*rt.at(rec, kRFields, 0, col));
  ro.Ip = T(*rt.at(rec, kRFields, 1, col));
  ro.Im = T(*rt.at(rec, kRFields, 2, col));
  ro.I1 = T(*rt.at(rec, kRFields, 3, col));
  ro.I2 = T(*rt.at(rec, kRFields, 4, col));
  return ro;
}

// Record of the segment starting at sample js (segments of `len` samples in
// the table's indexing), and of the last grid sample.
KG_INLINE int64_t segment_record(int64_t js, int64_t len) { return 2 + js / len; }
KG_INLINE int64_t last_record(int64_t ngeo, int64_t len) { return seg_count(ngeo, len) + 2; }

// Theta walk state from a direct evaluation at argument arg.
template <class T, int KIND, bool DO_PHI_T>
KG_INLINE void theta_direct(const ThetaConsts<T>& tc, const JacobiLadder<T>& lad, const WalkStep<T, 1>& w,
                            const T& arg, ThState<T>& ts, WalkState<T, 1>& st) {
  theta_state<T, KIND>(tc, lad, arg, DO_PHI_T, ts);
  st.s = ts.s;
  st.c = ts.c;
  st.d = ts.d;
  // k = 0 (|a| below MINSPIN): G_t needs no E (theta_out), and (E - F)/k is
  // not formed; the walked D stays finite and unused
  st.D = (DO_PHI_T && KIND != THK_ORD_K0) ? ts.EmF / tc.v[TH_k] : T(0.0);
  st.P[0] = T(0.0);
  st.Lg[0] = T(0.0);
  if (DO_PHI_T) walk_enter(w, 0, ts.Pi, arg, ts.s, ts.c, ts.d, st);
}

template <class T>
KG_INLINE void theta_init(const ThetaConsts<T>& tc, WalkStep<T, 1>& w) {
  const T n[1] = {tc.v[TH_n]};
  walk_init(w, tc.v[TH_k], n);
}

// Outputs of theta sample j from its state.
template <class T, int KIND, bool DO_PHI_T>
KG_INLINE bool theta_emit(const SampleOut& o, const ThetaConsts<T>& tc, int64_t i, int64_t j, double tau,
                          const T& arg, const ThState<T>& ts) {
  T th = T(0.0), Gph = T(0.0), Gt = T(0.0);
  theta_out<T, KIND>(tc, arg, ts, DO_PHI_T, th, Gph, Gt);
  return theta_store<T, DO_PHI_T>(o, j * o.npix + i, tau, th, Gph, Gt);
}

// Record store that marks a record the anchor pass could not reach (the
// walk pass then queues its segment for the fixup pass).
KG_INLINE void record_invalidate(const Records& rt, int64_t rec, int nf, int64_t col) {
  *rt.at(rec, nf, 0, col) = NAN;
}

// Heads of the anchor pass: 0 the step h, 1 the step L h, 2 the last grid
// sample, 2 + b (b >= 1) the first sample of block b - 1 (direct evaluation).
KG_INLINE int64_t anchor_heads(int64_t ngeo) { return 3 + anchor_blocks(ngeo); }

// Theta head hd of output pixel i (direct evaluation).
template <class T, int KIND, bool DO_PHI_T>
KG_INLINE bool theta_head(const SampleOut& o, const ThetaConsts<T>& tc, const JacobiLadder<T>& lad, int64_t i,
                          int64_t hd, double stop, const Records& rt, int64_t col) {
  constexpr int64_t L = kAnchorEveryLong;
  const int64_t last = o.ngeo - 1;
  WalkStep<T, 1> w;
  theta_init(tc, w);
  ThState<T> ts;
  WalkState<T, 1> st;
  if (hd < 2) {
    const double h = stop / static_cast<double>(o.ngeo - 1);
    const T arg = -tc.v[TH_w] * (hd == 0 ? h : static_cast<double>(L) * h);
    theta_direct<T, KIND, DO_PHI_T>(tc, lad, w, arg, ts, st);
    record_store<T, 1>(rt, hd, kThFields, col, st, static_cast<const T*>(nullptr));
    return false;
  }
  const int64_t j = (hd == 2) ? last : (hd - 3) * kAnchorBlock;
  if (hd > 2 && j >= last) return false;  // the last sample is head 2
  const double tau = mino_tau(j, o.ngeo, stop, false);
  const T arg = theta_arg(tc, T(tau));
  theta_direct<T, KIND, DO_PHI_T>(tc, lad, w, arg, ts, st);
  record_store<T, 1>(rt, hd == 2 ? last_record(o.ngeo, L) : segment_record(j, L), kThFields, col, st,
                     static_cast<const T*>(nullptr));
  return theta_emit<T, KIND, DO_PHI_T>(o, tc, i, j, tau, arg, ts);
}

// Theta segment starts of block blk after its first sample, each one step of
// L samples from the previous (addition theorems, no direct evaluation).  A
// refused step leaves the block's remaining records marked unreachable.
template <class T, int KIND, bool DO_PHI_T>
KG_INLINE bool theta_far(const SampleOut& o, const ThetaConsts<T>& tc, int64_t i, int64_t blk, double stop,
                         const Records& rt, int64_t col) {
  constexpr int64_t L = kAnchorEveryLong;
  const T m = tc.v[TH_k];
  const int64_t last = o.ngeo - 1;
  const int64_t jb = blk * kAnchorBlock;
  const int64_t je = (jb + kAnchorBlock < last) ? jb + kAnchorBlock : last;
  if (jb + L >= je) return false;
  WalkStep<T, 1> w;  // the long step
  theta_init(tc, w);
  record_load<T, 1>(rt, 1, kThFields, col, w.b, static_cast<T*>(nullptr));
  WalkState<T, 1> st;
  record_load<T, 1>(rt, segment_record(jb, L), kThFields, col, st, static_cast<T*>(nullptr));
  wait_vector_memory();
  bool bad = false;
  bool lost = false;
  for (int64_t js = jb + L; js < je; js += L) {
    WalkState<T, 1> nx;
    lost = lost || !walk_step<T, 1, DO_PHI_T, true>(st, w, nx);
    if (lost) {
      record_invalidate(rt, segment_record(js, L), kThFields, col);
      continue;
    }
    st = nx;
    record_store<T, 1>(rt, segment_record(js, L), kThFields, col, st, static_cast<const T*>(nullptr));
    const double tau = mino_tau(js, o.ngeo, stop, false);
    const T arg = theta_arg(tc, T(tau));
    ThState<T> ts;
    ts.s = st.s;
    ts.c = st.c;
    ts.d = st.d;
    ts.F = arg;
    ts.EmF = m * st.D;
    ts.Pi = DO_PHI_T ? walk_pi(w, 0, st, arg) : T(0.0);
    bad |= theta_emit<T, KIND, DO_PHI_T>(o, tc, i, js, tau, arg, ts);
  }
  return bad;
}

// Theta samples j0..j1-1 walked from the records.  Where a step is refused,
// `push(i, j, true)` queues samples j.. of the segment for the fixup pass,
// which evaluates each of them directly; the lane idles until the next
// segment (j stays uniform).
template <class T, int KIND, bool DO_PHI_T, class Push>
KG_INLINE bool theta_walk(const SampleOut& o, const ThetaConsts<T>& tc, int64_t i, int64_t j0, int64_t j1,
                          double stop, const Records& rt, int64_t col, Push& push) {
  constexpr int64_t L = kAnchorEveryLong;
  const T m = tc.v[TH_k];
  WalkStep<T, 1> w;
  theta_init(tc, w);
  record_load<T, 1>(rt, 0, kThFields, col, w.b, static_cast<T*>(nullptr));
  wait_vector_memory();
  const int64_t last = o.ngeo - 1;
  WalkState<T, 1> st;
  bool bad = false;
  bool queued = false;  // the rest of this segment is queued
  for (int64_t j = j0; j < j1; ++j) {
    if (j == last) continue;  // written by the anchor pass
    if (j % L == 0) {
      record_load<T, 1>(rt, segment_record(j, L), kThFields, col, st, static_cast<T*>(nullptr));
      wait_vector_memory();  // a queued segment would leave them pending into later steps
      queued = !kfinite(val(st.s));  // unreachable: the fixup pass takes the segment
      if (queued) push_range(push, i, j, j + L, last);
      continue;
    }
    if (queued) continue;
    WalkState<T, 1> nx;
    if (!walk_step<T, 1, DO_PHI_T, false, true>(st, w, nx)) {
      push_range(push, i, j, j - j % L + L, last);
      queued = true;
      continue;
    }
    st = nx;
    const int64_t idx = j * o.npix + i;
    const double tau = mino_tau(j, o.ngeo, stop, false);
    const T snarg = theta_arg(tc, T(tau));
    ThState<T> ts;
    ts.s = st.s;
    ts.c = st.c;
    ts.d = st.d;
    ts.F = snarg;
    ts.EmF = m * st.D;
    ts.Pi = DO_PHI_T ? walk_pi(w, 0, st, snarg) : T(0.0);
    T th = T(0.0), Gph = T(0.0), Gt = T(0.0);
    theta_out<T, KIND>(tc, snarg, ts, DO_PHI_T, th, Gph, Gt);
    bad |= theta_store<T, DO_PHI_T>(o, idx, tau, th, Gph, Gt);
  }
  return bad;
}

template <class T, int CASE>
KG_INLINE void radial_init(const RadialConsts<T>& rc, WalkStep<T, 3>& w) {
  T n[3], pc[3];
  radial_pi_params<T, CASE>(rc, n, pc);
  walk_init(w, rc.v[R_k], n);
}

// Radial walk state (and block logs) from a direct evaluation at argument X
// (dv: the Mino-time offset from the observer, or INFINITY for a plain step).
template <class T, int CASE, bool DO_PHI_T>
KG_INLINE void radial_direct(const RadialConsts<T>& rc, const JacobiLadder<T>& lad, const WalkStep<T, 3>& w,
                             const T& X, const T& dv, RState<T>& rs, WalkState<T, 3>& st, T (&lb)[3]) {
  constexpr bool kBlocks = DO_PHI_T && (CASE == 3 || CASE == 4);
  radial_state<T, CASE>(rc, lad, X, dv, DO_PHI_T, rs);
  st.s = rs.s;
  st.c = rs.c;
  st.d = rs.d;
  st.D = DO_PHI_T ? rs.EmF / rc.v[R_k] : T(0.0);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    st.P[q] = T(0.0);
    st.Lg[q] = T(0.0);
    if (DO_PHI_T) walk_enter(w, q, rs.Pi[q], X, rs.s, rs.c, rs.d, st);
    lb[q] = kBlocks ? block_log_enter(rs.lg[q]) : T(0.0);
  }
}

// Radial state for the outputs at argument X from a walked state.
template <class T>
KG_INLINE void radial_from_walk(const WalkStep<T, 3>& w, const WalkState<T, 3>& st, const T (&lb)[3], const T& X,
                                const T& m, bool do_phi_t, RState<T>& rs) {
  rs.s = st.s;
  rs.c = st.c;
  rs.d = st.d;
  rs.F = X;
  rs.EmF = m * st.D;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    rs.Pi[q] = do_phi_t ? walk_pi(w, q, st, X) : T(0.0);
    rs.lg[q] = lb[q];
  }
}

// Radial head hd of pixel column col (direct evaluation; heads as for
// theta).  Records 0 and 1 hold plain Jacobi functions of the step argument.
// The head writes records only: the radial walk emits every segment start
// from its record (radial_walk), so the anchor passes never read the theta
// pass's G_phi, G_t and can run while the theta passes do.
template <class T, int CASE, bool DO_PHI_T>
KG_INLINE void radial_head(const SampleOut& o, const RadialConsts<T>& rc, const JacobiLadder<T>& lad, int64_t hd,
                           double stop, const Records& rt, int64_t col) {
  constexpr int64_t L = (CASE == 2) ? kAnchorEveryLong : kAnchorEvery;
  const int64_t last = o.ngeo - 1;
  WalkStep<T, 3> w;
  radial_init<T, CASE>(rc, w);
  RState<T> rs;
  WalkState<T, 3> st;
  T lb[3];
  if (hd < 2) {
    const double h = stop / static_cast<double>(o.ngeo - 1);
    const T X = -rc.v[R_Xs] * (hd == 0 ? h : static_cast<double>(L) * h);
    radial_direct<T, CASE, DO_PHI_T>(rc, lad, w, X, T(INFINITY), rs, st, lb);
    lb[0] = lb[1] = lb[2] = T(0.0);  // no block logs at a step
    record_store<T, 3>(rt, hd, kRFields, col, st, lb);
    return;
  }
  const int64_t j = (hd == 2) ? last : (hd - 3) * kAnchorBlock;
  if (hd > 2 && j >= last) return;  // the last sample is head 2
  const double tau = mino_tau(j, o.ngeo, stop, false);
  const T dv = rc.v[R_Xs] * tau;
  const T X = rc.v[R_U0] - dv;
  radial_direct<T, CASE, DO_PHI_T>(rc, lad, w, X, dv, rs, st, lb);
  if (hd == 2) {
    // no step starts from the last sample: its record holds the sample's
    // radial outputs, from this direct evaluation, for the walk to emit
    last_out_store(rt, last_record(o.ngeo, kAnchorEvery), col, radial_out<T, CASE>(rc, T(tau), rs, DO_PHI_T));
    return;
  }
  record_store<T, 3>(rt, segment_record(j, kAnchorEvery), kRFields, col, st, lb);
}

// A far pass without a repair pass: a lost long step invalidates the
// block's remaining segment records, and the walk pass queues those segments
// whole for the fixup pass.
struct NoRepair {
  static constexpr bool kOn = false;
  KG_INLINE void operator()(int64_t, int64_t, bool) const {}
};

// Radial segment starts of block blk after its first sample (as theta_far),
// records only (the walk emits the samples).  With a repair pass
// (Repair::kOn), the first lost long step hands segment js and the rest of
// the block to it (repair(i, js, false)) instead.
template <class T, int CASE, bool DO_PHI_T, class Repair = NoRepair>
KG_INLINE void radial_far(const SampleOut& o, const RadialConsts<T>& rc, int64_t i, int64_t blk, const Records& rt,
                          int64_t col, const Repair& repair = Repair()) {
  constexpr int64_t L = (CASE == 2) ? kAnchorEveryLong : kAnchorEvery;
  constexpr bool kBlocks = DO_PHI_T && (CASE == 3 || CASE == 4);
  const int64_t last = o.ngeo - 1;
  const int64_t jb = blk * kAnchorBlock;
  const int64_t je = (jb + kAnchorBlock < last) ? jb + kAnchorBlock : last;
  if (jb + L >= je) return;
  WalkStep<T, 3> w;  // the long step
  radial_init<T, CASE>(rc, w);
  record_load<T, 3>(rt, 1, kRFields, col, w.b, static_cast<T*>(nullptr));
  WalkState<T, 3> st;
  T lb[3] = {T(0.0), T(0.0), T(0.0)};
  record_load<T, 3>(rt, segment_record(jb, kAnchorEvery), kRFields, col, st, kBlocks ? lb : static_cast<T*>(nullptr));
  wait_vector_memory();
  bool lost = false;
  for (int64_t js = jb + L; js < je; js += L) {
    WalkState<T, 3> nx;
    lost = lost || !walk_step<T, 3, DO_PHI_T, true>(st, w, nx) ||
           (kBlocks && !radial_block_step<T, CASE>(rc, st, nx, lb));
    if (lost) {
      if (Repair::kOn) {
        repair(i, js, false);
        return;
      }
      record_invalidate(rt, segment_record(js, kAnchorEvery), kRFields, col);
      continue;
    }
    st = nx;
    record_store<T, 3>(rt, segment_record(js, kAnchorEvery), kRFields, col, st, lb);
  }
}

// The repair pass: segment start js of a block whose far pass lost a long
// step at or before js, by direct evaluation; writes its record, as the far
// pass would have.
template <class T, int CASE, bool DO_PHI_T>
KG_INLINE void radial_repair(const SampleOut& o, const RadialConsts<T>& rc, const JacobiLadder<T>& lad, int64_t js,
                             double stop, const Records& rt, int64_t col) {
  WalkStep<T, 3> w;
  radial_init<T, CASE>(rc, w);
  const double tau = mino_tau(js, o.ngeo, stop, false);
  const T dv = rc.v[R_Xs] * tau;
  RState<T> rs;
  WalkState<T, 3> st;
  T lb[3];
  radial_direct<T, CASE, DO_PHI_T>(rc, lad, w, rc.v[R_U0] - dv, dv, rs, st, lb);
  record_store<T, 3>(rt, segment_record(js, kAnchorEvery), kRFields, col, st, lb);
}

// Segment starts of the block of a repair item (i, js0): js0 + s * L for
// s = 0, 1, ... while short of the block's end; -1 past it.
KG_INLINE int64_t repair_segment(int64_t js0, int64_t s, int64_t ngeo, int rcase) {
  const int64_t L = (rcase == 3 || rcase == 4) ? kAnchorEvery : kAnchorEveryLong;
  const int64_t last = ngeo - 1;
  const int64_t jb = js0 / kAnchorBlock * kAnchorBlock;
  const int64_t je = (jb + kAnchorBlock < last) ? jb + kAnchorBlock : last;
  const int64_t js = js0 + s * L;
  return js < je ? js : -1;
}

// Radial samples j0..j1-1 walked from the records.
template <class T, int CASE, bool DO_PHI_T, class Push>
KG_INLINE bool radial_walk(const SampleOut& o, const RadialConsts<T>& rc, const CombineConsts<T>& cmb, int64_t i,
                           int64_t j0, int64_t j1, double stop, const Records& rt, int64_t col, Push& push) {
  constexpr int64_t L = (CASE == 2) ? kAnchorEveryLong : kAnchorEvery;
  constexpr bool kBlocks = DO_PHI_T && (CASE == 3 || CASE == 4);
  const T m = rc.v[R_k];
  WalkStep<T, 3> w;
  radial_init<T, CASE>(rc, w);
  record_load<T, 3>(rt, 0, kRFields, col, w.b, static_cast<T*>(nullptr));
  wait_vector_memory();
  const int64_t last = o.ngeo - 1;
  WalkState<T, 3> st;
  T lb[3] = {T(0.0), T(0.0), T(0.0)};
  bool bad = false;
  bool queued = false;
  const int64_t jw = j1 < last ? j1 : last;
  for (int64_t j = j0; j < jw; ++j) {
    if (j % L == 0) {
      // a segment start: its state is the anchor passes' record, and the
      // sample is emitted from it here (the anchor passes write records
      // only).  A record that is unreachable, or whose state sits too close
      // to a singularity to be walked (non-finite entries), sends the
      // segment to the fixup pass, which evaluates it directly.
      record_load<T, 3>(rt, segment_record(j, kAnchorEvery), kRFields, col, st,
                        kBlocks ? lb : static_cast<T*>(nullptr));
      bool ok = kfinite(val(st.s)) & kfinite(val(st.c)) & kfinite(val(st.d)) & kfinite(val(st.D));
#pragma unroll
      for (int q = 0; q < 3; ++q)
        ok = ok & kfinite(val(st.P[q])) & kfinite(val(st.Lg[q])) & (!kBlocks || kfinite(val(lb[q])));
      queued = !ok;
      if (queued) {
        push_range(push, i, j, j + L, last);
        continue;
      }
      const int64_t idx = j * o.npix + i;
      const T gph = DO_PHI_T ? T(o.ph[idx]) : T(0.0);
      const T gt = DO_PHI_T ? T(o.t[idx]) : T(0.0);
      const double tau = mino_tau(j, o.ngeo, stop, false);
      RState<T> rs;
      radial_from_walk(w, st, lb, rc.v[R_U0] - rc.v[R_Xs] * tau, m, DO_PHI_T, rs);
      const RadialOut<T> ro = radial_out<T, CASE>(rc, T(tau), rs, DO_PHI_T);
      bad |= radial_store_g<T, DO_PHI_T>(o, cmb, idx, tau, ro, gph, gt);
      continue;
    }
    if (queued) continue;
    // the theta pass's G_phi, G_t of this sample, loaded before the step so
    // that their memory latency overlaps its arithmetic
    const int64_t idx = j * o.npix + i;
    const T gph = DO_PHI_T ? T(o.ph[idx]) : T(0.0);
    const T gt = DO_PHI_T ? T(o.t[idx]) : T(0.0);
    WalkState<T, 3> nx;
    if (!walk_step<T, 3, DO_PHI_T, false, true>(st, w, nx) ||
        (kBlocks && !radial_block_step<T, CASE, true>(rc, st, nx, lb))) {
      push_range(push, i, j, j - j % L + L, last);
      queued = true;
      continue;
    }
    st = nx;
    const double tau = mino_tau(j, o.ngeo, stop, false);
    RState<T> rs;
    radial_from_walk(w, st, lb, rc.v[R_U0] - rc.v[R_Xs] * tau, m, DO_PHI_T, rs);
    RadialOut<T> ro = radial_out<T, CASE>(rc, T(tau), rs, DO_PHI_T);
    bad |= radial_store_g<T, DO_PHI_T>(o, cmb, idx, tau, ro, gph, gt);
  }
  if (j1 > last && j0 <= last) {
    // the last sample: the head's direct outputs, kept in its record
    const int64_t idx = last * o.npix + i;
    const T gph = DO_PHI_T ? T(o.ph[idx]) : T(0.0);
    const T gt = DO_PHI_T ? T(o.t[idx]) : T(0.0);
    const RadialOut<T> ro = last_out_load<T>(rt, last_record(o.ngeo, kAnchorEvery), col);
    bad |= radial_store_g<T, DO_PHI_T>(o, cmb, idx, mino_tau(last, o.ngeo, stop, false), ro, gph, gt);
  }
  return bad;
}

// Per-pixel dispatch of the passes.  A theta pixel walks when its
// motion is ordinary or vortical; a radial pixel when its modulus is below 1.
struct ThetaPick {
  int kind;
  double stop;
  bool walk;
};

template <class T>
KG_INLINE ThetaPick theta_pick(const double* __restrict__ plan, int64_t stride, int64_t p, uint32_t gflags,
                               const SampleOut& o, ThetaConsts<T>& tc) {
#pragma unroll
  for (int s = 0; s < TH_NSLOT; ++s) tc.v[s] = T(plan[static_cast<int64_t>(PL_TH + s) * stride + p]);
  const int kind = static_cast<int>(plan[static_cast<int64_t>(PL_THK) * stride + p]);
  const bool eta0 = plan[static_cast<int64_t>(PL_ETA0) * stride + p] != 0.0;
  ThetaPick tp;
  tp.stop = plan[static_cast<int64_t>(PL_STOP) * stride + p];
  tp.kind = (eta0 && (gflags & GF_ANY_ETA_POS) == 0) ? THK_NONE : kind;
  tp.walk = walk_ok(o, (gflags & GF_ANY_STEP_ZERO) != 0) &&
            (tp.kind == THK_ORD || tp.kind == THK_ORD_K0 || tp.kind == THK_VORT);
  return tp;
}

// Theta head hd of output pixel i (plan column p).
template <bool DO_PHI_T, class T = double>
KG_INLINE uint32_t theta_head_pixel(const double* __restrict__ plan, int64_t stride, int64_t p, uint32_t gflags,
                                    const SampleOut& o, int64_t i, int64_t hd, const Records& rt) {
  ThetaConsts<T> tc;
  const ThetaPick tp = theta_pick(plan, stride, p, gflags, o, tc);
  if (!tp.walk) return 0u;
  JacobiLadder<T> lad;
  jacobi_ladder(tc.v[TH_mp], lad);
  bool bad;
  switch (tp.kind) {
    case THK_ORD: bad = theta_head<T, THK_ORD, DO_PHI_T>(o, tc, lad, i, hd, tp.stop, rt, p); break;
    case THK_ORD_K0: bad = theta_head<T, THK_ORD_K0, DO_PHI_T>(o, tc, lad, i, hd, tp.stop, rt, p); break;
    default: bad = theta_head<T, THK_VORT, DO_PHI_T>(o, tc, lad, i, hd, tp.stop, rt, p); break;
  }
  return loop_status(bad, lad);
}

// Theta segment starts of block blk of output pixel i after the block's head.
template <bool DO_PHI_T, class T = double>
KG_INLINE uint32_t theta_far_pixel(const double* __restrict__ plan, int64_t stride, int64_t p, uint32_t gflags,
                                   const SampleOut& o, int64_t i, int64_t blk, const Records& rt) {
  ThetaConsts<T> tc;
  const ThetaPick tp = theta_pick(plan, stride, p, gflags, o, tc);
  if (!tp.walk) return 0u;
  bool bad;
  switch (tp.kind) {
    case THK_ORD: bad = theta_far<T, THK_ORD, DO_PHI_T>(o, tc, i, blk, tp.stop, rt, p); break;
    case THK_ORD_K0: bad = theta_far<T, THK_ORD_K0, DO_PHI_T>(o, tc, i, blk, tp.stop, rt, p); break;
    default: bad = theta_far<T, THK_VORT, DO_PHI_T>(o, tc, i, blk, tp.stop, rt, p); break;
  }
  return bad ? ST_NONFINITE : 0u;
}

template <bool DO_PHI_T, class Push, class T = double>
KG_INLINE uint32_t theta_walk_pixel(const double* __restrict__ plan, int64_t stride, int64_t p, uint32_t gflags,
                                    const SampleOut& o, int64_t i, int64_t j0, int64_t j1, const Records& rt,
                                    Push& push) {
  ThetaConsts<T> tc;
  const ThetaPick tp = theta_pick(plan, stride, p, gflags, o, tc);
  if (!tp.walk) {
    for (int64_t js = j0; js < j1; js += kFixupSpan) push(i, js, false);
    return 0u;
  }
  bool bad;
  switch (tp.kind) {
    case THK_ORD: bad = theta_walk<T, THK_ORD, DO_PHI_T>(o, tc, i, j0, j1, tp.stop, rt, p, push); break;
    case THK_ORD_K0: bad = theta_walk<T, THK_ORD_K0, DO_PHI_T>(o, tc, i, j0, j1, tp.stop, rt, p, push); break;
    default: bad = theta_walk<T, THK_VORT, DO_PHI_T>(o, tc, i, j0, j1, tp.stop, rt, p, push); break;
  }
  return bad ? ST_NONFINITE : 0u;
}

// End of the fixup range that starts at sample j: the next multiple of
// kFixupSpan (which divides every segment length), short of the last grid
// sample when the anchor pass wrote it (walked: the radial store combines in
// place, so no sample may be stored twice).
KG_INLINE int64_t fixup_end(int64_t j, int64_t ngeo, bool walked) {
  const int64_t je = (j / kFixupSpan + 1) * kFixupSpan;
  const int64_t lim = walked ? ngeo - 1 : ngeo;
  return je < lim ? je : lim;
}

// Queued theta sample j + s (0 <= s < kFixupSpan) of the range j .. fixup_end,
// by direct evaluation, if it lies in the range: the fixup pass with one
// sample per lane, whose lanes are independent (the pass is latency-bound:
// few items, each a chain of dependent evaluations when walked).
template <bool DO_PHI_T, class T = double>
KG_INLINE uint32_t theta_fixup_sample(const double* __restrict__ plan, int64_t stride, int64_t p, uint32_t gflags,
                                      const SampleOut& o, int64_t i, int64_t j, bool walked, int64_t s) {
  const int64_t js = j + s;
  if (js >= fixup_end(j, o.ngeo, walked)) return 0u;
  return theta_pixel<DO_PHI_T, T, false>(plan, stride, p, gflags, o, i, js, js + 1);
}

template <class T>
KG_INLINE int radial_pick(const double* __restrict__ plan, int64_t stride, int64_t p, RadialConsts<T>& rc, T& lam,
                          double& stop) {
#pragma unroll
  for (int s = 0; s < R_NSLOT; ++s) rc.v[s] = T(plan[static_cast<int64_t>(PL_R + s) * stride + p]);
  rc.rcase = static_cast<int>(plan[static_cast<int64_t>(PL_RCASE) * stride + p]);
  lam = T(plan[static_cast<int64_t>(PL_LAM) * stride + p]);
  stop = plan[static_cast<int64_t>(PL_STOP) * stride + p];
  return (rc.rcase == 3 || rc.rcase == 4) ? kAnchorEvery : kAnchorEveryLong;
}

KG_INLINE bool radial_walks(const SampleOut& o, uint32_t gflags, double m) {
  return walk_ok(o, (gflags & GF_ANY_STEP_ZERO) != 0) && m < 1.0;
}

// A head (FAR = false: direct evaluation) or the long steps of a block
// (FAR = true) of one radial case, its slots loaded inside the case as in
// the walk.
// KIND: 0 a head (direct evaluation), 1 the long steps of a block, 2 the
// repair of a block from segment start k.
template <class T, int CASE, bool DO_PHI_T, int KIND, class Repair>
KG_INLINE uint32_t radial_anchor_case(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride,
                                      int64_t p, const SampleOut& o, int64_t i, int64_t k, const Records& rt,
                                      const Repair& repair) {
  RadialConsts<T> rc;
  T lam;
  double stop;
  radial_pick(plan, stride, p, rc, lam, stop);
  (void)c;
  if constexpr (KIND == 1) {
    radial_far<T, CASE, DO_PHI_T>(o, rc, i, k, rt, p, repair);
    return 0u;
  } else {
    JacobiLadder<T> lad;
    jacobi_ladder(rc.v[R_k], lad);
    if constexpr (KIND == 2) radial_repair<T, CASE, DO_PHI_T>(o, rc, lad, k, stop, rt, p);
    else radial_head<T, CASE, DO_PHI_T>(o, rc, lad, k, stop, rt, p);
    return loop_status(false, lad);
  }
}

template <bool DO_PHI_T, int KIND, class T = double, class Repair = NoRepair>
KG_INLINE uint32_t radial_anchor_dispatch(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride,
                                          int64_t p, uint32_t gflags, const SampleOut& o, int64_t i, int64_t k,
                                          const Records& rt, const Repair& repair = Repair()) {
  const int rcase = static_cast<int>(plan[static_cast<int64_t>(PL_RCASE) * stride + p]);
  const double m = plan[static_cast<int64_t>(PL_R + R_k) * stride + p];
  if (!radial_walks(o, gflags, m)) return 0u;
  switch (rcase) {
    case 4: return radial_anchor_case<T, 4, DO_PHI_T, KIND>(c, plan, stride, p, o, i, k, rt, repair);
    case 3: return radial_anchor_case<T, 3, DO_PHI_T, KIND>(c, plan, stride, p, o, i, k, rt, repair);
    default: return radial_anchor_case<T, 2, DO_PHI_T, KIND>(c, plan, stride, p, o, i, k, rt, repair);
  }
}

// Radial head hd of output pixel i (plan column p).
template <bool DO_PHI_T, class T = double>
KG_INLINE uint32_t radial_head_pixel(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride,
                                     int64_t p, uint32_t gflags, const SampleOut& o, int64_t i, int64_t hd,
                                     const Records& rt) {
  return radial_anchor_dispatch<DO_PHI_T, 0, T>(c, plan, stride, p, gflags, o, i, hd, rt);
}

// Radial segment starts of block blk of output pixel i after the block's head.
template <bool DO_PHI_T, class T = double, class Repair = NoRepair>
KG_INLINE uint32_t radial_far_pixel(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride,
                                    int64_t p, uint32_t gflags, const SampleOut& o, int64_t i, int64_t blk,
                                    const Records& rt, const Repair& repair = Repair()) {
  return radial_anchor_dispatch<DO_PHI_T, 1, T>(c, plan, stride, p, gflags, o, i, blk, rt, repair);
}

// Segment s of the repair item (i, js0) of output pixel i: the segment start
// js0 + s * L of js0's block, if short of the block's end (the repair pass
// for a block whose far pass lost a long step at js0).
template <bool DO_PHI_T, class T = double>
KG_INLINE uint32_t radial_repair_pixel(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride,
                                       int64_t p, uint32_t gflags, const SampleOut& o, int64_t i, int64_t js0,
                                       int64_t s, const Records& rt) {
  const int rcase = static_cast<int>(plan[static_cast<int64_t>(PL_RCASE) * stride + p]);
  const int64_t js = repair_segment(js0, s, o.ngeo, rcase);
  if (js < 0) return 0u;
  return radial_anchor_dispatch<DO_PHI_T, 2, T>(c, plan, stride, p, gflags, o, i, js, rt);
}

// The walk of one radial case: its plan slots are loaded here, inside the
// case, so the three cases' constants are never live together.
template <class T, int CASE, bool DO_PHI_T, class Push>
KG_INLINE bool radial_walk_case(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride, int64_t p,
                                const SampleOut& o, int64_t i, int64_t j0, int64_t j1, const Records& rt,
                                Push& push) {
  RadialConsts<T> rc;
  T lam;
  double stop;
  radial_pick(plan, stride, p, rc, lam, stop);
  const CombineConsts<T> cmb = combine_consts(c, lam);
  return radial_walk<T, CASE, DO_PHI_T>(o, rc, cmb, i, j0, j1, stop, rt, p, push);
}

template <bool DO_PHI_T, class Push, class T = double>
KG_INLINE uint32_t radial_walk_pixel(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride,
                                     int64_t p, uint32_t gflags, const SampleOut& o, int64_t i, int64_t j0,
                                     int64_t j1, const Records& rt, Push& push) {
  const int rcase = static_cast<int>(plan[static_cast<int64_t>(PL_RCASE) * stride + p]);
  const double m = plan[static_cast<int64_t>(PL_R + R_k) * stride + p];
  if (!radial_walks(o, gflags, m)) {
    for (int64_t js = j0; js < j1; js += kFixupSpan) push(i, js, false);
    return 0u;
  }
  bool bad;
  switch (rcase) {
    case 4: bad = radial_walk_case<T, 4, DO_PHI_T>(c, plan, stride, p, o, i, j0, j1, rt, push); break;
    case 3: bad = radial_walk_case<T, 3, DO_PHI_T>(c, plan, stride, p, o, i, j0, j1, rt, push); break;
    default: bad = radial_walk_case<T, 2, DO_PHI_T>(c, plan, stride, p, o, i, j0, j1, rt, push); break;
  }
  return bad ? ST_NONFINITE : 0u;
}

// Queued radial sample j + s of the range j .. fixup_end (as theta_fixup_sample).
template <bool DO_PHI_T, class T = double>
KG_INLINE uint32_t radial_fixup_sample(const RayScalars<T>& c, const double* __restrict__ plan, int64_t stride,
                                       int64_t p, uint32_t gflags, const SampleOut& o, int64_t i, int64_t j,
                                       bool walked, int64_t s) {
  const int64_t js = j + s;
  if (js >= fixup_end(j, o.ngeo, walked)) return 0u;
  return radial_pixel<DO_PHI_T, T, false>(c, plan, stride, p, gflags, o, i, js, js + 1);
}

// Per-pixel part of the plan kernels: the plan column, labels and the two
// image-wide predicates the reference's numpy code evaluates over all pixels.
struct SetupOut {
  bool eta_pos;
  bool step_zero;
};

// THETA_ONLY: the plan without its radial constants, which
// setup_radial_pixel then writes (the plan kernels' split).
template <bool THETA_ONLY = false>
KG_INLINE SetupOut setup_pixel(const RayScalars<double>& c, int64_t ngeo, bool do_phi_t, double al, double be,
                                       double* __restrict__ plan, int64_t stride, int64_t p, int8_t* rclass,
                                       int8_t* thclass, int32_t* nmax_eq, double* tau_tot, uint32_t* status,
                                       int64_t i) {
  double pl[PL_N];
  PixelSetup<double> px = THETA_ONLY ? plan_pixel_theta(c, al, be, do_phi_t, pl) : plan_pixel(c, al, be, do_phi_t, pl);
#pragma unroll
  for (int s = 0; s < PL_N; ++s)
    if (!THETA_ONLY || !plan_slot_radial(s)) plan[static_cast<int64_t>(s) * stride + p] = pl[s];
  SetupOut so;
  so.eta_pos = px.eta > 0.0;
  const double stop = pl[PL_STOP];
  so.step_zero = (ngeo > 1) && (stop / static_cast<double>(ngeo - 1) == 0.0);
  if (rclass) rclass[i] = static_cast<int8_t>(px.rclass);
  if (thclass) thclass[i] = static_cast<int8_t>(px.thclass);
  if (tau_tot) tau_tot[i] = px.tau_tot;
  if (nmax_eq) {
    // the last grid value is `stop` for ngeo > 1 in both linspace branches
    const double tl = (ngeo > 1) ? stop : 0.0 * stop;
    nmax_eq[i] = nmax_equatorial_at(c, al, be, tl);
  }
  if (status) status[i] = px.status;
  return so;
}

// The radial constants of pixel i's plan column p (the other half of
// setup_pixel<true>).
KG_INLINE void setup_radial_pixel(const RayScalars<double>& c, bool do_phi_t, double al, double be,
                                  double* __restrict__ plan, int64_t stride, int64_t p) {
  double pl[PL_N];
  plan_pixel_radial(c, al, be, do_phi_t, pl);
#pragma unroll
  for (int s = 0; s < PL_N; ++s)
    if (plan_slot_radial(s)) plan[static_cast<int64_t>(s) * stride + p] = pl[s];
}

#if defined(__HIPCC__)
__device__ inline void wave_or(uint32_t* word, bool pred, uint32_t bit) {
  const unsigned long long b = __ballot(pred ? 1 : 0);
  if (b != 0ull && __lane_id() == 0) atomicOr(word, bit);
}
#endif

}  // namespace kg
