/*
 * doma_oracle.h -- CPU restatement of DOMA's fixed-source discrete-ordinates
 * sweep and scattering source iteration (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP implementation and the timed CPU
 * baseline.  Only tests, the smoke check and the benchmark's CPU-baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every function follows the reference file:line quoted in its comment and
 * keeps the reference's floating-point expression order, so that with the same
 * OpenMP thread count it reproduces the reference bit for bit.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - quadrature: closed forms (numpy leggauss, Chebyshev nodes), weight sums,
 *     antipodal symmetry, and the reference's own build (oracle/_ref);
 *   - 3D/2D sweep + SI + MG: golden vectors written by the reference's own
 *     compiled TransportSolver2D/3D units (tests/golden/), plus the per-cell
 *     balance identity, DD linearity, TWDD(a=b=c=0.5) == DD, mirror symmetry;
 *   - 1D sweep + SI + MG: the reference's 1D solver needs Eigen, which is absent,
 *     so it cannot be built here; 1D is pinned by the exact DD recurrence in a
 *     pure absorber, the cell balance identity, mirror symmetry of symmetric
 *     decks and the infinite-medium limit.
 *   - 1D DSA (TransportSolver1D.cpp:587-800, needs Eigen, unbuildable here): the
 *     tridiagonal solve against numpy.linalg.solve; the accelerated iteration
 *     against the unaccelerated fixed point (the scheme is consistent) and its
 *     iteration counts against the unaccelerated ones.
 *   - fission source and power iteration: 2D/3D against the reference's own
 *     compiled solveFixedSource/solveEigenvalue on the fission decks
 *     (tests/golden/decks); 1D against the dominant eigenvalue of the assembled
 *     DD transport operator (numpy, tiny meshes) and the infinite-medium k.
 */
#ifndef DOMA_ORACLE_H
#define DOMA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_DD = 0, ORACLE_TW_DD = 1 };

typedef struct oracle_problem {
  int dims;                 /* 1, 2 or 3 */
  int nx, ny, nz;           /* cells per axis (ny = 1 in 1D, nz = 1 in 1D/2D) */
  const double *lx;         /* [nx] cell widths along x */
  const double *ly;         /* [ny] */
  const double *lz;         /* [nz] */
  int groups;
  int scheme;               /* ORACLE_DD or ORACLE_TW_DD */
  int num_mats;
  const int32_t *mat;       /* [nx*ny*nz] material id per cell, k-major, i fastest */
  const double *sigt;       /* [num_mats][groups] */
  const double *sigs;       /* [num_mats][groups][groups], [dest*G + src] */
  const double *src;        /* [num_mats][groups] */
  const uint8_t *has_src;   /* [num_mats] 0: no source vector (reference: empty _g_src) */
  int dirs;                 /* directions per octant (2D/3D) or per half range (1D) */
  const double *mu;         /* [num_oct][dirs] signed direction cosines, reference order */
  const double *eta;
  const double *xi;
  const double *w;          /* [num_oct][dirs] */
  /* Fission data (NULL chi: none).  chi/prod/prompt [num_mats][groups]: prompt
   * spectrum, nu*sigma_f, prompt fraction 1 - sum_d beta (1 without delayed
   * groups); has_chi/has_prod [num_mats]: 0 where the reference's vector is empty. */
  const double *chi;
  const double *prod;
  const double *prompt;
  const uint8_t *has_chi;
  const uint8_t *has_prod;
  int dsa;                  /* 1D: diffusion synthetic acceleration after every sweep */
  /* Transient data (oracle_solve_transient only; ORACLE_MAX_DELAYED slots per
   * material and group): inv_v [M][G]; nd [M]; chi_d, beta [M][G][8]; lambda
   * [M][8]; step_type [M] (-1 none, 0 both, 1 insert, 2 remove), step_in,
   * step_out [M], step_src [M][G] (NULL step_type: no step sources). */
  const double *inv_v;
  const int32_t *nd;
  const double *chi_d;
  const double *beta;
  const double *lambda;
  const int32_t *step_type;
  const double *step_in;
  const double *step_out;
  const double *step_src;
} oracle_problem;

#define ORACLE_MAX_DELAYED 8

typedef struct oracle_stats {
  int success;              /* reference solveFixedSource return value */
  int mg_iterations;
  double mg_residual;
  int sweeps;               /* total sweeps performed */
  int si_iterations_last[64];  /* SI count per group in the last MG pass */
  double si_residual_last[64];
  double k_eff;             /* oracle_solve_eigenvalue only */
  double k_residual;
} oracle_stats;

/* Quadrature (AngularPolynomials.cpp:8-111, GLAngularQuadrature.cpp:3-20,
 * GCAngularQuadrature.cpp:9-100). */
void oracle_legendre(int n, double *roots, double *weights);
void oracle_chebyshev(int n, double *roots, double *angular_roots, double *weights);
/* Fills mu/eta/xi/w as [8][n_c*n_l/4] (3D) or [4][n_c*n_l/4] (2D, octants with
 * xi > 0 in PPP, PMP, MPP, MMP order).  Returns directions per octant, or -1. */
int oracle_gc_quadrature(int n_c, int n_l, int dims, double *mu, double *eta, double *xi,
                         double *w);
/* 1D: mu/w as [2][n/2]: half 0 = mu > 0 (indices [0,n/2)), half 1 = mu < 0. */
int oracle_gl_quadrature(int n, double *mu, double *w);

/* Single-cell solves, for hand-checked values.
 * TWDiamondDifference3D.h:15-92; DiamondDifference3D.h:14-49;
 * TWDiamondDifference2D.h:15-76; DiamondDifference2D.h:14-45;
 * DiamondDifference1D.h:14-46.  in[] are upwind faces, out[] downwind faces;
 * returns the cell-centre angular flux. */
double oracle_twdd3d_cell(double q, double sigt, double lx, double ly, double lz,
                          double amu, double aeta, double axi, const double in[3],
                          double out[3]);
double oracle_dd3d_cell(double q, double sigt, double lx, double ly, double lz,
                        double amu, double aeta, double axi, const double in[3],
                        double out[3]);
double oracle_twdd2d_cell(double q, double sigt, double lx, double ly, double amu,
                          double aeta, double axi, const double in[2], double out[2]);
double oracle_dd2d_cell(double q, double sigt, double lx, double ly, double amu,
                        double aeta, const double in[2], double out[2]);
double oracle_dd1d_cell(double q, double sigt, double lx, double mu, double in,
                        double *out);

/* One full sweep of group g (TransportSolver3D.cpp:565-653 and 655-874,
 * TransportSolver2D.cpp:565-725, TransportSolver1D.cpp:840-915).
 * q[N] is the per-cell iteration source; dphi[N] receives the swept scalar
 * flux, reduced over per-thread partials in thread order. */
int oracle_sweep(const oracle_problem *p, int g, const double *q, double *dphi, int nthreads);

/* Full fixed-source solve (TransportSolver3D.cpp:262-563 and the 2D/1D
 * counterparts).  phi_out is [G][N].  Returns the reference's boolean result. */
int oracle_solve_fixed_source(const oracle_problem *p, double fs_tol, int fs_max_it,
                              double mg_tol, int mg_max_it, int nthreads, double *phi_out,
                              oracle_stats *stats);

/* k-eigenvalue power iteration (TransportSolver3D.cpp:29-86, 404-456 and the
 * 2D/1D counterparts).  Needs p->chi.  Writes k to *k_eff and stats. */
int oracle_solve_eigenvalue(const oracle_problem *p, double fs_tol, int fs_max_it,
                            double mg_tol, int mg_max_it, double k_tol, int nthreads,
                            double *phi_out, double *k_eff, oracle_stats *stats);

/* Transient from a zero initial condition (TransportSolver3D.cpp:88-260, ics =
 * zero; 2D/1D counterparts): dt = (t1 - t0) / steps, t from 0; per step the
 * multigroup solve with the time terms, then the precursor update.  phi_steps
 * [steps][G][N] and dnp_steps [steps][8][N] (may be NULL) receive every step's
 * flux and precursors; stats describe the last step solved.  Returns 1 if every
 * step converged. */
int oracle_solve_transient(const oracle_problem *p, double fs_tol, int fs_max_it, double mg_tol,
                           int mg_max_it, double t0, double t1, int steps, int nthreads,
                           double *phi_steps, double *dnp_steps, oracle_stats *stats);

/* Tridiagonal solve by Gaussian elimination with partial pivoting (row
 * interchanges between neighbours, as LAPACK dgtsv).  lo[i] = A[i][i-1],
 * di[i] = A[i][i], up[i] = A[i][i+1]; the arrays are overwritten.  Used by the
 * 1D DSA restatement; exported for its test. */
void oracle_tridiag_solve(int n, double *lo, double *di, double *up, double *rhs, double *x);

/* Alcouffe's DSA system of one 1D sweep (TransportSolver1D.cpp:587-790): phi,
 * cur (cell currents), sfr/sfl (right/left edge fluxes of each cell) and qmg are
 * [N]; lo/di/up/rhs [N+1] receive the tridiagonal rows.  Exported for tests. */
void oracle_dsa_system(const oracle_problem *p, int g, int eigen, const double *phi,
                       const double *cur, const double *sfr, const double *sfl,
                       const double *qmg, double *lo, double *di, double *up, double *rhs);

/* Source iteration of one group in isolation, starting from phi_tot[g] = 0 and
 * the given multigroup source qmg[N] (already scaled by the angular
 * normalisation).  Used to time sweeps and to trace residual sequences. */
int oracle_source_iteration(const oracle_problem *p, int g, const double *qmg,
                            double fs_tol, int fs_max_it, int nthreads, double *phi_g,
                            int *iterations, double *residual, double *residual_trace);

#ifdef __cplusplus
}
#endif
#endif
