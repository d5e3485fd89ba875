/*
 * kgeo C-ABI: MI355X (gfx950) FP64 analytic Kerr raytracer.
 *
 * Plain C types only.  Every array argument is a device pointer (HIP global
 * memory) owned by the caller; the library never allocates or frees caller
 * memory.  Kernels run asynchronously on the given HIP stream (NULL = the
 * default stream); a return of 0 means the work was enqueued.  Negative
 * returns are errors (see KGEO_E_*), with a message in kgeo_last_error()
 * (thread-local).  No C++ exception crosses this boundary.
 *
 * Layouts: per-pixel arrays are [npix]; per-sample arrays are [ngeo][npix]
 * (sample-major, pixels contiguous), the layout of the reference's (ngeo,
 * npix) numpy arrays.  Output pointers marked "optional" may be NULL.
 *
 * Numerical problems never fail a call: they are reported per pixel in the
 * optional status bitmask (KGEO_ST_*), and the affected outputs carry
 * NaN / inf exactly where the reference's numpy code produces them.
 */
#ifndef KGEO_H_
#define KGEO_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KGEO_ABI_VERSION 4

/* return codes */
#define KGEO_OK 0
#define KGEO_E_ARG (-1)       /* invalid argument (size, pointer, range) */
#define KGEO_E_HIP (-2)       /* HIP runtime / launch error */
#define KGEO_E_WORKSPACE (-3) /* workspace too small or misaligned */

/* params.flags */
#define KGEO_DO_PHI_T 1u /* sample t, phi, sigma: required by kgeo_plan, kgeo_sample,
                            kgeo_raytrace_ana, kgeo_coords_at_tau and every image of
                            kgeo_batch_plan (the reference's raytrace_ana always samples
                            them, kerr_raytracing_ana.py:97); calls without it return
                            KGEO_E_ARG */

/* per-pixel status bits */
#define KGEO_ST_NONFINITE 1u         /* some output sample is NaN or inf */
#define KGEO_ST_DEGENERATE_ROOTS 2u  /* resolvent root xi0 <= 0 (B = 0) */
#define KGEO_ST_MINSPIN_VORTICAL 4u  /* |a| < 1e-4 and eta <= 0: the reference raises */
#define KGEO_ST_LADDER_CAP 8u        /* Jacobi AGM ladder capped */
#define KGEO_ST_LABEL_TIE 16u        /* rclass 0 (exact tie r4 == r_+) or cube-root tie */

/*
 * Ray-wide parameters.  a, r_o, th_o, maxtaufrac are the reference's
 * arguments (kerr_raytracing_ana.py:46-51).  The derived scalars must be
 * the values numpy gives for the reference's expressions:
 *   a2 = a**2, sin_th = np.sin(th_o), cos_th = np.cos(th_o),
 *   cos2_th = np.cos(th_o)**2, rplus = 1 + np.sqrt(1 - a**2),
 *   rminus = 1 - np.sqrt(1 - a**2)
 * so that the conserved quantities, and hence the integer labels, come out
 * bit-identical.  kgeo_fill_params() fills them with the C library's libm.
 */
typedef struct kgeo_ray_params {
  double a, r_o, th_o, maxtaufrac;
  double a2, sin_th, cos_th, cos2_th, rplus, rminus;
  int64_t npix, ngeo;
  uint32_t flags;
  uint32_t reserved;
} kgeo_ray_params;

int kgeo_abi_version(void);
const char* kgeo_last_error(void);

/* Fill the derived scalars of p from p->a and p->th_o (host-side, libm). */
void kgeo_fill_params(kgeo_ray_params* p);

/* Bytes of device workspace for npix pixels sampled on a linspace grid of
 * ngeo Mino times: 256 bytes of flags and counters, the per-pixel plan and,
 * for ngeo > 2, the walk's record tables and fixup queue (about 5.1 kB per
 * pixel at ngeo = 500).  Pass ngeo = 0 for kgeo_plan alone and for
 * caller-given Mino times (kgeo_coords_at_tau, kgeo_sample with tau_in).
 * Environment: KGEO_CHUNKS=<n> fixes the sample chunks per pixel. */
size_t kgeo_workspace_bytes(int64_t npix, int64_t ngeo);

/*
 * Replaces kerr_raytracing_ana.raytrace_ana (kerr_raytracing_ana.py:46-122):
 * conserved quantities, radial roots and class, total Mino time, the
 * linspace(0, maxtaufrac*tau_tot, ngeo) Mino-time grid and the coordinates
 * at every grid point (coords_at_tau, :126-193).
 *   alpha, beta      [npix]        screen coordinates
 *   tau              [ngeo][npix]  Geodesics.mino_times
 *   t, r, th, ph     [ngeo][npix]  Geodesics.geo_coords (t and ph are required:
 *                                   the theta pass parks G_phi, G_t in them)
 *   sig              [ngeo][npix]  Geodesics.affine_times
 *   (tau, r, th and sig may be NULL to skip that output)
 *   rclass           [npix] int8   optional: radial_roots rclass (utils.py:366-378)
 *   thclass          [npix] int8   optional: angular_turning thclass (utils.py:266-272)
 *   nmax_eq          [npix] int32  optional: Geodesics.nmax_eq (utils.py:63-67)
 *   tau_tot          [npix]        optional: mino_total (utils.py:382-548)
 *   status           [npix] uint32 optional: KGEO_ST_* bits
 *   workspace        >= kgeo_workspace_bytes(npix, ngeo) bytes, 256-byte aligned
 */
int kgeo_raytrace_ana(const kgeo_ray_params* p, const double* alpha, const double* beta, double* tau, double* t,
                      double* r, double* th, double* ph, double* sig, int8_t* rclass, int8_t* thclass,
                      int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace,
                      size_t workspace_bytes, void* stream);

/*
 * The two stages of kgeo_raytrace_ana, for callers that re-sample one plan.
 * kgeo_plan is the per-pixel preliminaries of raytrace_ana
 * (kerr_raytracing_ana.py:75-94: lambda, eta, radial_roots, mino_total) plus
 * the tau-independent parts of th_integrate / r_integrate (:217-251,
 * :300-327, :404-434, :473-510, :554-584), written into `workspace`.
 * kgeo_sample is coords_at_tau (:97, :126-193) from that plan: tau_in
 * [ntau][npix] if non-NULL, else the linspace grid of p->ngeo samples
 * (then ntau must equal p->ngeo, and the workspace must hold
 * kgeo_workspace_bytes(npix, ntau)).  Outputs as in kgeo_raytrace_ana.
 */
int kgeo_plan(const kgeo_ray_params* p, const double* alpha, const double* beta, int8_t* rclass, int8_t* thclass,
              int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace, size_t workspace_bytes,
              void* stream);
int kgeo_sample(const kgeo_ray_params* p, const double* tau_in, int64_t ntau, double* tau, double* t, double* r,
                double* th, double* ph, double* sig, uint32_t* status, void* workspace,
                size_t workspace_bytes, void* stream);

/*
 * No counterpart in the reference.  The number of Mino-time chunks per pixel
 * that kgeo_sample / kgeo_raytrace_ana split ngeo into for npix pixels (the
 * launch shape; results do not depend on it).  For measurement tools.
 */
int kgeo_sample_chunks(int64_t npix, int64_t ngeo);

/*
 * Replaces kerr_raytracing_ana.coords_at_tau (kerr_raytracing_ana.py:126-193):
 * coordinates at caller-given Mino times tau_in [ntau][npix].  Outputs as in
 * kgeo_raytrace_ana with ngeo = ntau (p->ngeo is ignored).
 */
int kgeo_coords_at_tau(const kgeo_ray_params* p, const double* alpha, const double* beta, const double* tau_in,
                       int64_t ntau, double* t, double* r, double* th, double* ph, double* sig, uint32_t* status,
                       void* workspace, size_t workspace_bytes, void* stream);

/*
 * Replaces n_equatorial_crossings (kerr_raytracing_utils.py:605-684) and
 * n_poloidal_orbits (kerr_raytracing_utils.py:551-603) at Mino times
 * tau [ntau][npix]: floor-valued equatorial crossings (-2 for eta <= 0) and
 * fractional poloidal orbits (-2 for eta <= 0), both [ntau][npix] float64.
 * Either output may be NULL.  Uses eta without the spin-zero fix, as the
 * reference does.
 */
int kgeo_crossings(const kgeo_ray_params* p, const double* alpha, const double* beta, const double* tau,
                   int64_t ntau, double* n_equatorial, double* n_poloidal, void* stream);

/*
 * Replaces n_angular_turnings (kerr_raytracing_utils.py:686-792) and adds
 * the radial turning count the reference decides inline
 * (equatorial_images.py:266-288), at Mino times tau [ntau][npix]:
 *   n_angular   [ntau][npix] float64  n_angular_turnings as the reference
 *                                     returns it (fractional; eta <= 0 rays
 *                                     take its vortical formula, :753-786)
 *   turns_theta [ntau][npix] int32    floor(n_angular): theta turning points
 *                                     passed by tau (INT32_MIN where n_angular
 *                                     is NaN, i.e. eta == 0)
 *   turns_r     [ntau][npix] int8     radial turning points passed by tau: 1
 *                                     where the ray is outside the critical
 *                                     curve (is_outside_crit, :794-826) and
 *                                     tau > tau_tot / 2 (mino_total, :382-548,
 *                                     with raytrace_ana's spin-zero eta fix), else 0
 * Any output may be NULL.  n_angular and turns_theta need th_o in [0, pi/2],
 * as the reference's n_angular_turnings does (:690-691; KGEO_E_ARG otherwise).
 */
int kgeo_turnings(const kgeo_ray_params* p, const double* alpha, const double* beta, const double* tau,
                  int64_t ntau, double* n_angular, int32_t* turns_theta, int8_t* turns_r, void* stream);

/*
 * Replaces the elliptic plug-in ellippi_arr(n, phi, m)
 * (scipy_ellip_binding.py:87-237; GSL twin gsl_ellip_binding.py:154-237):
 * elementwise Pi(n; phi | m) in the A&S convention with the same special
 * cases (NaN / inf / limits) and the Cauchy principal value for n > 1.
 * All arrays [count].
 */
int kgeo_ellip_pi(const double* n, const double* phi, const double* m, double* out, int64_t count, void* stream);

/*
 * Equatorial-emission image (workload composed of equatorial_lensing.py
 * r_equatorial :85-196 and equatorial_images.py :175, :266-334 with the
 * velocities.py u_kep :229-281 disk): for subrings mbar = 0..mbar_max (<= 3),
 * I_mbar = g^3 r_s^-emis_p where the subring emits (not N_max < mbar,
 * N_max == -1 or r_s <= r_+), image = sum_mbar I_mbar.
 *   image [npix]; I_sub, g_sub, r_s_sub, Ir_sub [mbar_max+1][npix] (r_s and
 *   I_r as r_equatorial returns them); Imax [npix]; Nmax [npix] int32 (-2 for
 *   eta <= 0); zero_mask [npix] uint8, bit mbar set where subring mbar is
 *   masked.  Every output may be NULL.  Uses p->a, th_o, r_o and the
 *   derived scalars; p->ngeo and p->maxtaufrac are ignored.
 */
int kgeo_equatorial_image(const kgeo_ray_params* p, const double* alpha, const double* beta, int mbar_max,
                          double emis_p, double* image, double* I_sub, double* g_sub, double* r_s_sub,
                          double* Ir_sub, double* Imax, int32_t* Nmax, uint8_t* zero_mask, void* stream);

/*
 * No counterpart in the reference (it has no gradients).  Vector-Jacobian
 * product of kgeo_equatorial_image with respect to (a, th_o, emis_p):
 * grad3[k] = sum_i grad_image[i] * d image[i] / d theta_k, each pixel's
 * tangent computed by recomputing its geodesic with forward-mode dual
 * numbers (no sample storage), reduced deterministically (fixed order, no
 * atomics).  grad_image may be NULL (all ones).  jac [3][npix] (optional)
 * receives the per-pixel tangents.  workspace >=
 * kgeo_equatorial_vjp_workspace_bytes(npix) device bytes.
 */
size_t kgeo_equatorial_vjp_workspace_bytes(int64_t npix);
int kgeo_equatorial_image_vjp(const kgeo_ray_params* p, const double* alpha, const double* beta, int mbar_max,
                              double emis_p, const double* grad_image, double* jac, double* grad3, void* workspace,
                              size_t workspace_bytes, void* stream);
const char* kgeo_equatorial_last_error(void);

/*
 * Replaces rho_of_req / rho_of_req_single (equatorial_lensing.py:226-262)
 * and the root finds of generate_library (:351-416) for many curves in one
 * launch.  Curve k is the lensed image of the equatorial circle r = req[k]
 * in subring mbar[k] for the observer curves[k] (a, th_o and the derived
 * scalars as for kgeo_raytrace_ana; the reference places the observer at
 * r_o = infinity, :19, :202).  For each screen angle varphi[i]:
 *   rho[k][i]    the screen radius at which r_equatorial2 (:198-216: r_s of
 *                the mbar-th crossing; 1e50 for rays outside the critical
 *                curve without that crossing, 1 for rays inside it) equals
 *                req[k], by Brent's method over the reference's bracket
 *                (mbar 0: [0, req + 5]; 1: [0, 100]; else [1, 10]) with
 *                brentq's tolerances (xtol 2e-12, rtol 4 eps, 100 iterations)
 *   status[k][i] optional int32: 0 found; 1 no sign change over the bracket
 *                (brentq raises ValueError there); 2 not converged
 *                (RuntimeError); rho is NaN unless 0
 * All arrays are device arrays, curves included ([ncurves] kgeo_ray_params).
 */
int kgeo_rho_of_req(const kgeo_ray_params* curves, const double* req, const int32_t* mbar, int64_t ncurves,
                    const double* varphi, int64_t nphi, double* rho, int32_t* status, void* stream);

/*
 * Replaces the closed form of critical_curve (equatorial_lensing.py:264-293)
 * at photon-orbit radii r [n]: alpha [n], beta [n] >= 0 of the bound orbit at
 * each radius for p->a, p->th_o (NaN where it is not seen from th_o).  The
 * caller builds r = linspace(r_-, r_+, n) and drops the NaN points and
 * mirrors beta as the reference does (kgeo_amd.lensing.critical_curve).
 */
int kgeo_critical_curve(const kgeo_ray_params* p, const double* r, int64_t n, double* alpha, double* beta,
                        void* stream);

/*
 * Emission models of the equatorial disk: the reference's Velocity
 * (velocities.py:23-176), Bfield (bfields.py:33-427) and Emissivity
 * (emissivities.py:37-118) objects as plain descriptors.  Each field is the
 * reference's constructor argument of the same name, with its default
 * filled in by the caller (kgeo_amd.models does so).  Tables are device
 * arrays of 4 x n doubles, column blocks [r | c1 | c2 | c3] with r
 * increasing, interpolated as np.interp does.
 */
#define KGEO_VEL_ZAMO 0
#define KGEO_VEL_INFALL 1
#define KGEO_VEL_KEP 2        /* 'kep', 'cunningham' */
#define KGEO_VEL_SUBKEP 3     /* 'subkep', 'cunningham_subkep' */
#define KGEO_VEL_GENERAL 4
#define KGEO_VEL_GELLES 5
#define KGEO_VEL_SIMFIT 6
#define KGEO_VEL_FROMFILE 7   /* table: r, U1, U2, U3 */
#define KGEO_VEL_DRIFTFRAME 8 /* numeric nu_parallel, no gammamax */

#define KGEO_BF_SIMPLE 0      /* 'rad', 'vert', 'tor', 'simple' */
#define KGEO_BF_SIMPLE_RM1 1
#define KGEO_BF_MONOPOLE_A 2
#define KGEO_BF_BZ_MONOPOLE 3
#define KGEO_BF_BZ_GUESS 4
#define KGEO_BF_BZ_PARA 5
#define KGEO_BF_POWER 6
#define KGEO_BF_GEN_POWER 7
#define KGEO_BF_FROMFILE 8    /* table: r, Br, Btheta, Bphi */

#define KGEO_EMIS_CONSTANT 0
#define KGEO_EMIS_BPL 1
#define KGEO_EMIS_GLM 2       /* 'glm' and 'ring' (mu_ring > 0) */
#define KGEO_EMIS_THERMAL 3

typedef struct kgeo_bfield_model {
  int32_t kind;             /* KGEO_BF_* */
  int32_t secondorder_only; /* bz_monopole */
  int32_t usemono;          /* power */
  int32_t is_above;         /* gen_power isAbove */
  double C;                 /* bz_monopole, bz_guess, bz_para, monopoleA, power */
  double Cr, Cvert, Cph;    /* simple, simple_rm1 */
  double omegafac;          /* monopoleA */
  double p;                 /* power */
  double n_I, p_val;        /* gen_power */
  const double* table;      /* fromfile */
  int64_t table_n;
} kgeo_bfield_model;

typedef struct kgeo_velocity_model {
  int32_t kind;             /* KGEO_VEL_* */
  int32_t retrograde;       /* kep, subkep, general */
  double fac_subkep, beta_phi, beta_r;  /* subkep, general */
  double gelles_beta, gelles_chi;       /* gelles */
  double ell_isco, vr_isco, p1, p2, dd; /* simfit */
  double nu_parallel;                   /* driftframe */
  kgeo_bfield_model drift_field;        /* driftframe: the field it drifts in */
  const double* table;                  /* fromfile */
  int64_t table_n;
} kgeo_velocity_model;

typedef struct kgeo_emissivity_model {
  int32_t kind;             /* KGEO_EMIS_* */
  int32_t bfield_model;     /* thermal: take |B| from the field model */
  double emiscut_in, emiscut_out;
  double p1, p2;            /* bpl */
  double mu_ring;           /* glm: 0 means 1 - sqrt(1 - a^2); ring: r_ring */
  double gamma_off, sigma;  /* glm, ring */
  double Rb, ne0, Te0, B0, alpha_n, alpha_T, alpha_B; /* thermal */
} kgeo_emissivity_model;

typedef struct kgeo_emission_model {
  kgeo_velocity_model velocity;
  kgeo_bfield_model bfield;
  kgeo_emissivity_model emissivity;
  int32_t polarization, pathlength;     /* Iobs flags */
  double specind, diskangle, nu_obs;
} kgeo_emission_model;

/*
 * Replaces make_image / Iobs (equatorial_images.py:23-264) with
 * calc_redshift (:307-334), calc_polquantities (:336-487), calc_evpa
 * (:524-544) and calc_pathlength_equatorial (:489-522), for subrings
 * mbar = 0..mbar_max in one launch (one lane per pixel).  Emission is
 * evaluated at theta = pi/2 as numpy rounds it.  Outputs are [npix][nsub]
 * (nsub = mbar_max + 1, the reference's (npix, mbar_max+1) arrays):
 *   I, Q, U      observed Stokes intensities (0 where masked; Q = U = 0
 *                without polarization)
 *   g, sin_thb   redshift and sin(theta_B) (0 where masked; sin_thb is pi/4
 *                without polarization, as the reference assumes)
 *   r_s, Ir      emission radius and Mino time as r_equatorial returns them
 *   lp, Ie       with pathlength: path length and rest-frame emission
 * and per pixel: nmax_eq (N_max of r_equatorial, -2 for eta <= 0) and
 * n_poloidal (nmax_poloidal: n_poloidal_orbits at the total Mino time).
 * Every output may be NULL.  r_o must be the reference's (>= 100 or inf).
 * Models the reference cannot image return KGEO_E_ARG: a comoving field,
 * pathlength with a velocity that has u^theta != 0 (driftframe).
 */
int kgeo_equatorial_stokes(const kgeo_ray_params* p, const double* alpha, const double* beta, int32_t mbar_max,
                           const kgeo_emission_model* model, double* I, double* Q, double* U, double* g,
                           double* r_s, double* sin_thb, double* Ir, double* nmax_eq, double* n_poloidal,
                           double* lp, double* Ie, void* stream);

/*
 * The models at given points, replacing Velocity.u_lab (velocities.py:70-96),
 * Velocity.tetrades (:124-176), Bfield.bfield_lab / omega_field / bsq
 * (bfields.py:93-131, :175-196, :365-395) and Emissivity.jrest
 * (emissivities.py:83-118).  For point i at radius r[i] and polar angle
 * th[i] (th NULL: pi/2 as numpy rounds it):
 *   u [4][n]      contravariant four-velocity (model->velocity)
 *   tetrad [16][n] e_(a)^mu, rows (t, x, y, z) x columns mu (velocity)
 *   B [3][n]      lab-frame B^r, B^theta, B^phi (model->bfield)
 *   omega [n]     field-line angular velocity (model->bfield)
 *   bsq [n]       fluid-frame b^2 of model->bfield in model->velocity
 *   j [n]         jrest(a, r, g, sinthetab, Bmag, nu_obs, specind) of
 *                 model->emissivity; g, sinthb, bmag [n] may be NULL (None)
 * Every output may be NULL.  Models whose quantity the reference does not
 * define (omega of a simple field) return KGEO_E_ARG.
 */
int kgeo_model_eval(const kgeo_emission_model* model, double a, const double* r, const double* th, int64_t n,
                    const double* g, const double* sinthb, const double* bmag, double* u, double* tetrad, double* B,
                    double* omega, double* bsq, double* j, void* stream);
const char* kgeo_emission_last_error(void);

/*
 * No counterpart in the reference.  kgeo_equatorial_image_vjp without a
 * workspace, for callers that cannot allocate device memory (code under HIP
 * graph capture, e.g. a captured optimiser step over (a, th_o, p); tight
 * loops; harnesses that forbid hipMalloc).  It calls no hipMalloc / hipFree:
 * the per-block partial sums live in a static device array inside the
 * library, one copy per GPU (768 KiB, held until the process ends), whose
 * address is looked up once per device on the first call there.  Results
 * are bit-identical to kgeo_equatorial_image_vjp.  npix is limited to
 * KGEO_VJP_NOALLOC_MAX_NPIX (KGEO_E_ARG beyond).  Every call on one GPU
 * shares that GPU's array: calls that overlap in time on one GPU (two
 * streams, two threads) race and must be serialised by the caller.
 */
#define KGEO_VJP_NOALLOC_MAX_NPIX 8388608
int kgeo_equatorial_image_vjp_noalloc(const kgeo_ray_params* p, const double* alpha, const double* beta,
                                      int mbar_max, double emis_p, const double* grad_image, double* jac,
                                      double* grad3, void* stream);

/*
 * No counterpart in the reference (its callers loop over raytrace_ana, one
 * image per call).  Batched survey: nimg images, each with its own
 * kgeo_ray_params (a, r_o, th_o, npix, ngeo, flags), traced by one
 * persistent launch captured in a HIP graph.  Each image's results equal a
 * kgeo_raytrace_ana call on that image alone (same kernels' code): the
 * image-wide predicates (any eta > 0, any zero grid step) are per image.
 *
 * Layout (kgeo_batch_layout gives the offsets, nimg + 1 entries each):
 *   alpha, beta, rclass, thclass, nmax_eq, tau_tot, status: concatenated
 *     per-pixel arrays, image k at pix_off[k];
 *   t, r, th, ph: one arena, image k's [ngeo_k][npix_k] block at out_off[k].
 * tau and sigma are not stored (tau is the image's linspace grid; sigma can
 * be recomputed with kgeo_coords_at_tau).  Every image must set
 * KGEO_DO_PHI_T and t, ph are required; rclass..status are optional.
 *
 * kgeo_batch_plan uploads the image descriptors, records the graph on
 * `stream` (must not be NULL) and returns a handle; it runs no kernels.
 * kgeo_batch_run replays the whole survey on any stream.  The handle keeps
 * the pointers it was given: buffers and workspace must outlive it, and runs
 * that overlap in time on one handle race.  kgeo_batch_progress copies the
 * per-image completed work-item counts of the last run to the host (after
 * the run's stream is synchronised), so a caller can tell finished images
 * from unfinished ones.
 */
typedef struct kgeo_batch_tuning {
  int32_t samples_per_item; /* Mino-time samples per work item (rounded up to a multiple of 16); an item is 64 pixels x this */
  int32_t waves_per_cu;     /* persistent grid: CUs x this waves (1..64) */
  int32_t reserved[2];
} kgeo_batch_tuning;

typedef struct kgeo_batch kgeo_batch;

size_t kgeo_batch_workspace_bytes(const kgeo_ray_params* imgs, int32_t nimg);
int kgeo_batch_layout(const kgeo_ray_params* imgs, int32_t nimg, int64_t* pix_off, int64_t* out_off);
int kgeo_batch_plan(const kgeo_ray_params* imgs, int32_t nimg, const kgeo_batch_tuning* tune, const double* alpha,
                    const double* beta, double* t, double* r, double* th, double* ph, int8_t* rclass,
                    int8_t* thclass, int32_t* nmax_eq, double* tau_tot, uint32_t* status, void* workspace,
                    size_t workspace_bytes, void* stream, kgeo_batch** out);
int kgeo_batch_run(kgeo_batch* b, void* stream);
int64_t kgeo_batch_items(const kgeo_batch* b);
int kgeo_batch_progress(const kgeo_batch* b, uint32_t* done_items, uint32_t* total_items);
int kgeo_batch_destroy(kgeo_batch* b);
const char* kgeo_batch_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* KGEO_H_ */
