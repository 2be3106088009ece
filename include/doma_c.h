/*
 * doma_c.h -- C ABI of the MI355X-native discrete-ordinates sweep and
 * scattering source iteration.
 *
 * Every entry point returns an int status (DOMA_OK on success) and never exits
 * the process; doma_last_error() gives the message of the last failure on the
 * calling thread.  No C++ or GPU-runtime types cross this boundary: meshes,
 * materials and quadratures are plain arrays, host pointers are borrowed for
 * the duration of a call, and device memory is owned by the context.  Calls on
 * one context are serialised on the context's HIP stream; use one context per
 * host thread.
 *
 * Reference seams replaced (paths relative to the nuclearkevin/doma checkout):
 *   - T::solve policy + TransportSolver3D<T>::sweep (include/TransportSolver3D.h:
 *     15-16, 76; src/TransportSolver3D.cpp:565-874; include/TWDiamondDifference3D.h:
 *     15-92; include/DiamondDifference3D.h:14-49)      -> doma_sweep
 *   - TransportSolver3D<T>::sourceIteration (src/TransportSolver3D.cpp:458-563)
 *                                                       -> doma_source_iteration
 *   - TransportSolver3D<T>::solveFixedSource (include/TransportSolver3D.h:28;
 *     src/TransportSolver3D.cpp:262-402, 499-529)       -> doma_solve_fixed_source
 *   - TransportSolver3D<T>::solveEigenvalue (src/TransportSolver3D.cpp:29-86,
 *     404-456)                                          -> doma_solve_eigenvalue
 *   - TransportSolver3D<T>::solveTransient (src/TransportSolver3D.cpp:88-260)
 *                                  -> doma_transient_begin / doma_transient_step
 *   - the 2D and 1D counterparts (src/TransportSolver2D.cpp:262-725,
 *     src/TransportSolver1D.cpp:292-586, 802-915)       -> same entry points, dims 2/1
 */
#ifndef DOMA_C_H
#define DOMA_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DOMA_OK 0
#define DOMA_ERR_INVALID 1   /* bad argument or input deck semantics (reference: std::exit(1)) */
#define DOMA_ERR_DEVICE 2    /* HIP runtime failure or missing device */
#define DOMA_ERR_STATE 3     /* call order violated (e.g. solve before set_materials) */

#define DOMA_SCHEME_DD 0     /* DiamondDifference{1,2,3}D */
#define DOMA_SCHEME_TW_DD 1  /* TWDiamondDifference{2,3}D, theta = theta_s = 0.9 */

#define DOMA_FP64 0          /* everything in FP64 */
#define DOMA_FP32_SWEEP 1    /* FP32 angular flux + cell solve; FP64 scalar-flux
                                accumulation, sources and residuals */

#define DOMA_MAX_GROUPS 64
#define DOMA_MAX_DELAYED 8   /* delayed-neutron precursor groups per material */

#define DOMA_STEP_NONE (-1)  /* step source types (SourceStep, include/InputParameters.h:10-21) */
#define DOMA_STEP_BOTH 0     /* on for insert_at <= t <= remove_at */
#define DOMA_STEP_INSERT 1   /* on for insert_at <= t */
#define DOMA_STEP_REMOVE 2   /* on for t <= remove_at */

typedef struct doma_ctx doma_ctx;

/* Problem geometry, discretisation and angular quadrature.
 * Cells are ordered k-major, i fastest: c = (k*ny + j)*nx + i, as
 * BrickMesh3D (src/BrickMesh3D.cpp:86-102).  In 2D nz = 1, in 1D ny = nz = 1.
 * Directions are [num_octants][dirs] in the reference's order: 8 octants
 * PPP..MMM in 3D (GCAngularQuadrature(2na, 2np, 3)), 4 quadrants PPP, PMP,
 * MPP, MMP in 2D (GCAngularQuadrature(2na, 2np, 2)), 2 half ranges (mu > 0,
 * mu < 0) in 1D (GLAngularQuadrature(2np)).  Signed cosines; the sweeps use
 * their magnitudes as the reference does. */
typedef struct doma_problem_desc {
  int32_t dims;
  int32_t nx, ny, nz;
  const double *lx, *ly, *lz;  /* [nx], [ny], [nz] cell widths (cm) */
  int32_t groups;              /* 1 .. DOMA_MAX_GROUPS */
  int32_t scheme;              /* DOMA_SCHEME_DD | DOMA_SCHEME_TW_DD (TW_DD: dims 2, 3) */
  int32_t precision;           /* DOMA_FP64 | DOMA_FP32_SWEEP */
  int32_t dirs;                /* directions per octant / quadrant / half range */
  const double *mu, *eta, *xi, *w;
} doma_problem_desc;

/* Kernel shape of the sweep.  Zero fields mean "any" (the library default, or
 * the autotuned choice once doma_autotune has run for this problem shape);
 * non-zero fields restrict the choice.  doma_get_tuning reports the shape in use
 * with every field filled in. */
typedef struct doma_tuning {
  int32_t brick;          /* 3D: brick edge in cells (4 or 8); 2D: tile edge (8 or 16) */
  int32_t threads;        /* work-group size of the sweep kernel */
  int32_t exact_arith;    /* 1: reference expression order with IEEE divides,
                             0: hoisted reciprocals (default) */
  int32_t schedule;       /* 0: one launch per brick diagonal (default in 3D);
                             1: one persistent dataflow launch per sweep;
                             2: 2D only, one lane per mesh row, one work-group per
                                direction (default in 2D up to 512 rows);
                             3: 3D only, persistent dataflow launch whose tasks
                                are pencils of bricks along z */
  int32_t octants;        /* octants (2D: quadrants) per work-group: 1, 2, 4 or 8 */
  int32_t dirs_per_lane;  /* directions per lane: 1 or 2 */
  int32_t shape;          /* explicit kernel shape id + 1 (overrides the hints) */
} doma_tuning;

typedef struct doma_stats {
  int32_t success;               /* reference solveFixedSource / sourceIteration result */
  int32_t mg_iterations;
  double mg_residual;
  int32_t sweeps;                /* group sweeps performed */
  int32_t si_iterations[DOMA_MAX_GROUPS];  /* SI count per group, last MG pass */
  double si_residual[DOMA_MAX_GROUPS];
  double elapsed_ms;             /* wall time of the call on the context stream */
  double sweep_ms;               /* time spent in sweep kernels (HIP events on
                                    the first two sweeps of each group's source
                                    iteration; its later sweeps are counted at
                                    the second one's duration) */
  int32_t corrections_accepted;  /* learned corrections kept (doma_set_learned_correction) */
  int32_t corrections_rejected;  /* learned corrections undone because the residual rose */
  double correction_ms;          /* wall time spent evaluating the correction network */
  double k_eff;                  /* doma_solve_eigenvalue: final eigenvalue (else 0) */
  double k_residual;             /* doma_solve_eigenvalue: |k - k_prev| of the last pass */
} doma_stats;

const char *doma_last_error(void);
const char *doma_version(void);

/* Create a context on the current HIP device and upload geometry/quadrature. */
int doma_create(const doma_problem_desc *desc, doma_ctx **out);
int doma_destroy(doma_ctx *ctx);

/* Material map and tables (MaterialProps, include/InputParameters.h:24-55):
 * mat[N] ids in [0, num_mats); sigt[M][G]; sigs[M][G][G] dest-major
 * ([g_dest*G + g_src], TransportSolver3D.cpp:375); src[M][G]; has_src[M] = 0
 * for a material without an external source (reference: empty _g_src). */
int doma_set_materials(doma_ctx *ctx, const int32_t *mat, int32_t num_mats, const double *sigt,
                       const double *sigs, const double *src, const uint8_t *has_src);

/* Fission data of the materials set by doma_set_materials (MaterialProps
 * _g_chi_p, _g_prod, _g_n_beta, include/InputParameters.h:24-55): chi_p[M][G]
 * prompt fission spectrum, prod[M][G] nu*sigma_f, prompt[M][G] the prompt
 * fraction 1 - sum_d beta[g][d] of a material with delayed groups (1 without),
 * has_chi[M] / has_prod[M] = 0 where the reference's vector is empty.  A
 * material with chi_p must have nu_sigma_f.  Fixed-source solves add the prompt
 * fission source to the multigroup source (TransportSolver3D.cpp:377-387,
 * TransportSolver2D.cpp:377-387, TransportSolver1D.cpp:410-420);
 * doma_solve_eigenvalue requires this call.  doma_set_materials clears it. */
int doma_set_fission(doma_ctx *ctx, int32_t num_mats, const double *chi_p, const double *prod,
                     const double *prompt, const uint8_t *has_chi, const uint8_t *has_prod);

/* 1D diffusion synthetic acceleration (TransportSolver1D.cpp:587-800, the
 * reference's dsa="true"): after every sweep the flux is replaced by the solution
 * of Alcouffe's diffusion system (source correction in fixed-source solves,
 * modified diffusion coefficients or removal correction in eigenvalue solves),
 * solved on the device.  1D contexts only; rejects void cross sections (< 1e-6,
 * InputParameters.cpp:403-414).  enable = 0 turns it off. */
int doma_set_dsa(doma_ctx *ctx, int32_t enable);

/* Where source iteration tests its stopping rule (no reference counterpart):
 * device_loop = 1 (default) evaluates !(tol < r) and the iteration limit on the
 * device and lets the host enqueue several iterations per round trip (kernels
 * after the last iteration return at entry); 0 waits for the residual after
 * every sweep, as the reference's loop does.  Both give the same iterates and
 * counts.  The learned correction always uses the per-iteration loop. */
int doma_set_si_mode(doma_ctx *ctx, int32_t device_loop);

int doma_set_tuning(doma_ctx *ctx, const doma_tuning *t);
int doma_get_tuning(const doma_ctx *ctx, doma_tuning *t);

/* One full sweep of group g over all octants: dphi[c] = sum_n w_n psi_n(c) for the
 * per-cell iteration source q[c] (already scaled by the angular normalisation,
 * e.g. 0.25/pi in 3D).  Host arrays of N doubles. */
int doma_sweep(doma_ctx *ctx, int32_t g, const double *q, double *dphi);

/* Same, on device arrays of N doubles, enqueued on the context stream. */
int doma_sweep_device(doma_ctx *ctx, int32_t g, const double *d_q, double *d_dphi);

/* Scattering source iteration of one group from phi_g = 0 with multigroup source
 * qmg[N] (host).  Writes phi_g[N] (host, may be NULL), the iteration count and
 * last residual; *converged is the reference's sourceIteration result. */
int doma_source_iteration(doma_ctx *ctx, int32_t g, const double *qmg, double fs_tol,
                          int32_t fs_max_it, double *phi_g, int32_t *iterations,
                          double *residual, int32_t *converged);

/* Full fixed-source solve: Gauss-Seidel over groups, each group solved by
 * source iteration from zero (solveFixedSource semantics).  phi[G][N] (host) may
 * be NULL; the flux stays resident and can be read with doma_get_flux. */
int doma_solve_fixed_source(doma_ctx *ctx, double fs_tol, int32_t fs_max_it, double mg_tol,
                            int32_t mg_max_it, double *phi, doma_stats *stats);

/* k-eigenvalue power iteration (TransportSolver3D.cpp:29-86, 404-456;
 * TransportSolver2D.cpp:29-86, 404-456; TransportSolver1D.cpp:36-90, 437-493):
 * phi = 1, k = 1; each pass forms every group's in-scatter + fission/k source
 * and runs its source iteration (Gauss-Seidel over groups), then updates
 * k = sum(V nuSf phi) / sum(V nuSf phi_prev / k) (1D: k_prev * sum(l nuSf phi) /
 * sum(l nuSf phi_prev)); stops after mg_max_it passes or once the flux shape
 * residual <= mg_tol and |k - k_prev| <= k_tol.  A source iteration that does
 * not converge ends the solve unsuccessfully, as in the reference.  The
 * eigenvalue is returned in *k_eff (may be NULL) and stats->k_eff. */
int doma_solve_eigenvalue(doma_ctx *ctx, double fs_tol, int32_t fs_max_it, double mg_tol,
                          int32_t mg_max_it, double k_tol, double *phi, double *k_eff,
                          doma_stats *stats);

/* Kinetics data for transient solves (MaterialProps _g_inv_v, _num_d_groups,
 * _n_g_chi_d, _g_n_beta, _n_lambda, include/InputParameters.h:24-55): inv_v[M][G];
 * num_delayed[M] in [0, DOMA_MAX_DELAYED]; chi_d[M][G][DOMA_MAX_DELAYED],
 * beta[M][G][DOMA_MAX_DELAYED], lambda[M][DOMA_MAX_DELAYED] (entries past a
 * material's count are ignored).  A material with delayed groups needs
 * nu_sigma_f (doma_set_fission); the prompt fraction given there should be
 * 1 - sum_d beta.  doma_set_materials clears it. */
int doma_set_kinetics(doma_ctx *ctx, int32_t num_mats, const double *inv_v,
                      const int32_t *num_delayed, const double *chi_d, const double *beta,
                      const double *lambda);

/* Step sources of a transient (the deck's <transients><source_step>,
 * InputParameters.cpp:419-500), per material: type[M] (DOMA_STEP_*),
 * insert_at[M], remove_at[M], src[M][G].  doma_set_materials clears them. */
int doma_set_step_sources(doma_ctx *ctx, int32_t num_mats, const int32_t *type,
                          const double *insert_at, const double *remove_at, const double *src);

/* Transient from a zero initial condition (solveTransient with ics="zero",
 * TransportSolver3D.cpp:88-260 and the 2D/1D counterparts): begin sets phi,
 * the previous-step flux and the precursors to zero, dt = (t1 - t0) / steps and
 * t = 0 (the reference's step clock starts at zero).  Each doma_transient_step
 * solves one step as a fixed-source problem at time t with the implicit-Euler
 * terms (sigma_t + inv_v/dt in the sweep; inv_v/dt phi_last, the delayed source
 * chi_d lambda C and the step sources in the multigroup source), then advances
 * the precursors C_d = S_d/lambda (1 - e^{-lambda dt}) + C_d e^{-lambda dt} and t.
 * *t_solved (may be NULL) receives the step's time; the step's flux is read with
 * doma_get_flux and the precursors with doma_get_dnps. */
int doma_transient_begin(doma_ctx *ctx, double t0, double t1, int32_t steps);
int doma_transient_step(doma_ctx *ctx, double fs_tol, int32_t fs_max_it, double mg_tol,
                        int32_t mg_max_it, double *t_solved, doma_stats *stats);

/* Precursor concentrations [nd_max][N] (nd_max = largest num_delayed; zero where
 * a cell's material has fewer groups, as BrickMesh3D::dumpDNPsToTextFile). */
int doma_get_dnps(doma_ctx *ctx, double *dnps, int32_t *nd_max);

int doma_get_flux(doma_ctx *ctx, double *phi);

/* Time n back-to-back sweeps of group g on the resident source (HIP events on
 * the context stream); returns milliseconds per sweep and per-launch figures. */
int doma_time_sweeps(doma_ctx *ctx, int32_t g, int32_t n, double *ms_per_sweep,
                     int32_t *launches_per_sweep);

/* Profile n sweeps of group g with a HIP event pair around every kernel launch on
 * the context stream: kernel_ms = summed launch durations per sweep, wall_ms =
 * sweep wall time on the stream, launches = kernel launches per sweep. */
int doma_profile_sweeps(doma_ctx *ctx, int32_t g, int32_t n, double *kernel_ms, double *wall_ms,
                        int32_t *launches);

/* Number of cell-direction units one sweep of one group processes. */
int64_t doma_units_per_sweep(const doma_ctx *ctx);

/* Project-written autotuner: times the candidate kernel shapes and schedules
 * on this context's problem (each sweep followed by the sum of its
 * direction-set partials, as in an iteration) and keeps the fastest (cached per
 * problem shape); *best_ms is that per-iteration time. */
int doma_autotune(doma_ctx *ctx, int32_t g, doma_tuning *best, double *best_ms);

/* ------------------------------------------------------------------------- */
/* Learned source-iteration correction (no reference counterpart; replaces the  */
/* reference's hand-built acceleration, TransportSolver1D.cpp:587-800, in 3D). */
/* ------------------------------------------------------------------------- */

/* Number of FP32 parameters of the correction network (a 3-level 3D U-Net,
 * channels 12/24/32): per layer W[cout][tap][cin] then bias[cout]. */
int64_t doma_net_param_count(void);

/* Enable the correction on a 3D context (extents divisible by 4): every `every`
 * source iterations (every >= 2) the remaining error of the group is estimated
 * as g (d_k + s y): the geometric-series extrapolation g d_k of the latest
 * plain increment, g = rho / (1 - rho), rho = rms(d_k) / rms(d_(k-1)) clamped
 * to 0.999, s = rms(d_k), refined by the network's output y (BF16 on MFMA, FP32
 * accumulate), which reads d_k / s, d_(k-1) / s, sigma_t and the in-group
 * scattering ratio.  A zero network gives the bare extrapolation.  The
 * correction is undone if the next residual is not below the one before it.
 * every = 0 disables it; params may then be NULL. */
int doma_set_learned_correction(doma_ctx *ctx, const float *params, int64_t count, int32_t every);

/* Accelerated multigroup iteration (no reference counterpart; off by default):
 * from the second multigroup pass on, each group's source iteration starts from
 * the previous pass's flux instead of zero (the reference restarts from zero,
 * TransportSolver3D.cpp:399-400).  2D/3D. */
int doma_set_warm_start(doma_ctx *ctx, int32_t enable);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): doma_net_forward on buffers the caller owns, allocating
 * nothing.  d_params (the flat FP32 parameters, `count` values), d_features
 * ([V][4] FP32, V = D*H*W, x fastest) and d_out ([V] FP32) are device pointers;
 * the work is enqueued on `stream` (a hipStream_t; NULL = the null stream) and
 * the call returns without waiting.  The packed BF16 weights and every
 * activation live in one static device array of the library (under 8 MiB per
 * device, held until the process ends), so D, H, W must be multiples of 4 with
 * D*H*W <= DOMA_NET_NOALLOC_MAX_VOXELS (32^3), else DOMA_ERR_INVALID.  Calls
 * share that array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller; calls in order on one
 * stream are safe (each clears what it needs first). */
#define DOMA_NET_NOALLOC_MAX_VOXELS 32768
int doma_net_forward_noalloc(const float *d_params, int64_t count, int32_t D, int32_t H,
                             int32_t W, const float *d_features, float *d_out, void *stream);

/* Stopping rule of source iteration (no reference counterpart for rule 1).
 * DOMA_STOP_RESIDUAL (default): the reference's rule, stop once the relative
 * increment r_k = ||dphi_k|| / ||phi_k|| <= tol (TransportSolver3D.cpp:458-497).
 * DOMA_STOP_ERROR_ESTIMATE: stop once r_k * rho / (1 - rho) <= tol, where rho is
 * the largest ratio r_k / r_(k-1) of consecutive plain source iterations (no
 * correction in between) seen so far in the solve for the group (clamped to
 * 0.9999) -- the geometric-series estimate of the remaining error
 * phi_converged - phi_k relative to phi_k; the group keeps its estimate across
 * multigroup passes.  Rule 1 runs the iteration with one host round trip per
 * sweep. */
#define DOMA_STOP_RESIDUAL 0
#define DOMA_STOP_ERROR_ESTIMATE 1
int doma_set_stopping_rule(doma_ctx *ctx, int32_t rule);

/* Stand-alone forward pass (diagnostics and tests): features[V][4] (FP32, V =
 * D*H*W, x fastest) -> out[V]. */
int doma_net_forward(const float *params, int64_t count, int32_t D, int32_t H, int32_t W,
                     const float *features, float *out);

/* Hand-written training of the correction network on batches of `batch` cubic
 * crops of edge `crop` (multiple of 4): BF16 MFMA forward and data gradients,
 * FP32 weight gradients, loss sum (p - y)^2 / sum y^2, global-norm clipping at
 * 1, Adam (0.9, 0.999, 1e-8) on FP32 master weights.  features are
 * [batch][crop^3][4] and targets [batch][crop^3] (x fastest), on the host. */
typedef struct doma_net_trainer doma_net_trainer;
int doma_net_trainer_create(int32_t crop, int32_t batch, const float *params, int64_t count,
                            doma_net_trainer **out);
int doma_net_trainer_destroy(doma_net_trainer *t);
int doma_net_trainer_step(doma_net_trainer *t, const float *features, const float *targets,
                          float lr, float *loss);
/* Loss and gradient [count] without an update (tests). */
int doma_net_trainer_gradient(doma_net_trainer *t, const float *features, const float *targets,
                              float *loss, float *grad);
int doma_net_trainer_get_params(doma_net_trainer *t, float *params);

/* ------------------------------------------------------------------------- */
/* Batches: many independent 3D problems, one persistent sweep launch per      */
/* source-iteration round (no reference counterpart).                          */
/* ------------------------------------------------------------------------- */

typedef struct doma_batch doma_batch;

typedef struct doma_batch_info {
  int32_t rounds;      /* batched sweep launches (source-iteration rounds) */
  int64_t units;       /* cell-direction-group updates swept, all problems */
  double elapsed_ms;   /* wall time of doma_batch_solve */
  double sweep_ms;     /* device time of the batched sweep launches */
  int64_t samples;     /* training snapshots kept */
} doma_batch_info;

/* count problems; all 3D, same scheme, precision and angular quadrature (any
 * extents, any groups).  One shared stream; the sweep kernel shape is common to
 * the batch. */
int doma_batch_create(int32_t count, const doma_problem_desc *descs, doma_batch **out);
int doma_batch_destroy(doma_batch *b);
int doma_batch_set_materials(doma_batch *b, int32_t p, const int32_t *mat, int32_t num_mats,
                             const double *sigt, const double *sigs, const double *src,
                             const uint8_t *has_src);
/* Kernel shape for the batch (schedule is always the persistent dataflow one). */
int doma_batch_set_tuning(doma_batch *b, const doma_tuning *t);
int doma_batch_get_tuning(const doma_batch *b, doma_tuning *t);
/* Times every valid 3D kernel shape on one batched sweep of all problems and
 * keeps the fastest. */
int doma_batch_autotune(doma_batch *b, doma_tuning *best, double *best_ms);
/* n batched sweeps of group 0 of every problem: device ms per launch, units per launch. */
int doma_batch_time_sweeps(doma_batch *b, int32_t n, double *ms_per_round, int64_t *units);
/* Every problem solved like doma_solve_fixed_source.  With sample_every > 0,
 * (residual, converged - current flux) pairs are kept every sample_every
 * iterations of the first multigroup pass, at most sample_cap per (problem,
 * group), thinned evenly.  stats: [count] or NULL. */
int doma_batch_solve(doma_batch *b, double fs_tol, int32_t fs_max_it, double mg_tol,
                     int32_t mg_max_it, int32_t sample_every, int32_t sample_cap,
                     doma_stats *stats, doma_batch_info *info);
int doma_batch_get_flux(doma_batch *b, int32_t p, double *phi);
/* Snapshots kept for (p, g) and their iteration numbers (iterations may be NULL). */
int doma_batch_sample_count(const doma_batch *b, int32_t p, int32_t g, int32_t *count,
                            int32_t *iterations);
/* Snapshot k of (p, g): residual[N] = the latest SI increment d_k, prev_residual[N]
 * = the increment before it d_(k-1), target[N] = phi_conv - phi_k (any may be NULL). */
int doma_batch_get_sample(doma_batch *b, int32_t p, int32_t g, int32_t k, float *residual,
                          float *prev_residual, float *target);

/* ------------------------------------------------------------------------- */
/* Input decks (host C++ front end).                                           */
/* ------------------------------------------------------------------------- */

typedef struct doma_deck_info {
  int32_t dims, nx, ny, nz;      /* cells per axis after interval expansion and refinement */
  int32_t groups, scheme, mode;  /* mode: 0 steady (fixed source), 1 eigen, 2 transient */
  int32_t np, na;                /* polar / azimuthal quadrature parameters */
  int32_t num_materials, dsa, refinement;
  int32_t fs_max_it, mg_max_it;
  double fs_tol, mg_tol;
  double total_volume;
} doma_deck_info;

/* Read a <doma_inp> deck (src/InputParameters.cpp:91-524) and expand its mesh
 * (src/BrickMesh3D.cpp:8-197).  Host only; invalid decks give DOMA_ERR_INVALID
 * with the reference's message in doma_last_error(). */
int doma_deck_read(const char *path, doma_deck_info *out);

/* Run a fixed-source (solve_type="steady"), eigenvalue ("eigen") or transient
 * ("transient") deck end to end like the reference's main() (src/main.cpp:51-298): mesh,
 * solver by dimension and scheme, GPU solve; with a non-NULL output_base the
 * reference's text files <base>_dims.txt, _blocks.txt, _meshx/y/z.txt,
 * _g<g>_flux.txt are written.  phi (may be NULL) receives [G][N] of capacity
 * phi_len doubles; *converged is the reference's solveFixedSource /
 * solveEigenvalue / solveTransient result and stats->k_eff the eigenvalue.  A
 * transient deck (ics="zero") writes <base>_t<step>_g<g>_flux.txt and
 * _d<d>_dnps.txt per step; phi and stats are then the last step's.  tuning may be
 * NULL; autotune != 0 runs doma_autotune first. */
int doma_run_deck(const char *path, const char *output_base, int32_t precision,
                  const doma_tuning *tuning, int32_t autotune, double *phi, int64_t phi_len,
                  doma_stats *stats, int32_t *converged);

/* ------------------------------------------------------------------------- */
/* Angular quadratures (host only).                                            */
/* ------------------------------------------------------------------------- */

/* GCAngularQuadrature(n_c, n_l, dims) (src/GCAngularQuadrature.cpp:9-55): writes
 * [8][n_c*n_l/4] (3D) or [4][n_c*n_l/4] (2D) signed directions and weights in the
 * layout of doma_problem_desc.  The reference builds it with n_c = 2*na and
 * n_l = 2*np (src/TransportSolver3D.cpp:10).  Returns directions per octant, or
 * -1 (zero direction component for odd n_c / n_l, or a bad argument). */
int doma_gc_quadrature(int32_t n_c, int32_t n_l, int32_t dims, double *mu, double *eta, double *xi,
                       double *w);

/* GLAngularQuadrature(n) (src/GLAngularQuadrature.cpp:3-20): n Legendre nodes,
 * descending, so [0, n/2) is mu > 0.  Returns n/2, or -1. */
int doma_gl_quadrature(int32_t n, double *mu, double *w);

/* ------------------------------------------------------------------------- */
/* Synthetic seeded media (counter-based SplitMix64, identical on every host).  */
/* ------------------------------------------------------------------------- */

#define DOMA_SCATTER_DOWN1 0  /* sigma_s[g][g] = 0.7 c sigma_t, [g+1][g] = 0.3 c sigma_t;
                                 the last group keeps only its in-group part */
#define DOMA_SCATTER_FULL 1   /* in-group plus down-scatter to every lower group,
                                 up-scatter from the last group into the one above */
#define DOMA_SCATTER_INGROUP 2 /* sigma_s[g][g] = 0.98 c sigma_t, [g+1][g] = 0.02 c sigma_t;
                                 the last group keeps all of c in-group (down-scatter
                                 only: the in-group iterations dominate) */
#define DOMA_SOURCE_CENTRAL 0 /* unit source in group 0 on the cube [src_lo, src_hi)^3 */
#define DOMA_SOURCE_BLOCKS 1  /* Q_g ~ U[0,1] in a seeded fraction of blocks */

typedef struct doma_media_spec {
  int32_t n;               /* cells per side of the cubic mesh */
  int32_t block;           /* cells per side of a material block (divides n) */
  int32_t groups;
  int32_t sigt_log;        /* 0: sigma_t ~ U[lo, hi]; 1: log-uniform */
  double sigt_lo, sigt_hi;
  double c_lo, c_hi;       /* scattering ratio c ~ U[lo, hi] per (block, group) */
  int32_t scatter_rule;
  int32_t source_rule;
  int32_t src_lo, src_hi;  /* DOMA_SOURCE_CENTRAL */
  double source_fraction;  /* DOMA_SOURCE_BLOCKS */
  double cell_width;       /* cm */
  uint64_t seed;
} doma_media_spec;

/* Number of materials the spec produces (2 * blocks for DOMA_SOURCE_CENTRAL,
 * blocks otherwise). */
int32_t doma_media_num_materials(const doma_media_spec *spec);

/* Fills mat[n^3], sigt[M][G], sigs[M][G][G], src[M][G], has_src[M]. Host only. */
int doma_media_generate(const doma_media_spec *spec, int32_t *mat, double *sigt, double *sigs,
                        double *src, uint8_t *has_src);

#ifdef __cplusplus
}
#endif
#endif /* DOMA_C_H */
