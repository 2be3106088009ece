/*
 * doma_oracle.c -- CPU restatement of DOMA's fixed-source S_N sweep and
 * scattering source iteration.  TEST INFRASTRUCTURE: the checker for the HIP
 * path and the timed CPU baseline; never linked into the product.
 *
 * Expression order follows the reference line by line (cited per function), so
 * that with the same OpenMP team size the results equal the reference's bit for
 * bit.  Interface fluxes use compact per-thread buffers (one x face, one row of
 * y faces, one slice of z faces) instead of the reference's six faces per cell:
 * every upwind face a cell reads is the one its upwind neighbour wrote earlier
 * in the same direction's sweep, or zero on a vacuum boundary, exactly as
 * CartesianCell3D::interfaceFlux / boundaryFlux return (CartesianCell3D.cpp:31-38,
 * 93-103).
 */
#include "doma_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* Quadrature                                                                 */
/* ------------------------------------------------------------------------- */

/* LegendrePolynomial::evaluate (AngularPolynomials.cpp:45-66). */
static double legendre_eval(int n, double x)
{
  if (n == 0) return 1.0;
  if (n == 1) return x;
  double v = 1.0, v1 = x, v2 = 1.0;
  for (int i = 2; i <= n; ++i) {
    v = ((2.0 * (double)i - 1.0) * x * v1 - ((double)i - 1.0) * v2) / (double)i;
    v2 = v1;
    v1 = v;
  }
  return v;
}

/* LegendrePolynomial::evaluateDerivative (AngularPolynomials.cpp:68-94). */
static double legendre_deriv(int n, double x)
{
  if (n == 0) return 0.0;
  if (n == 1) return 1.0;
  double v = 1.0, d = 0.0, v1 = x, v2 = 1.0;
  double f = 1.0 / (x * x - 1.0);
  for (int i = 2; i <= n; ++i) {
    v = ((2.0 * (double)i - 1.0) * x * v1 - ((double)i - 1.0) * v2) / (double)i;
    d = (double)i * f * (x * v - v1);
    v2 = v1;
    v1 = v;
  }
  return d;
}

/* LegendrePolynomial ctor: Newton iteration from cos(pi(i-1/4)/(n+1/2)),
 * stop at |dr| <= 2e-16, w = 2/((1-x^2) P'^2); roots descending
 * (AngularPolynomials.cpp:8-39). */
void oracle_legendre(int n, double *roots, double *weights)
{
  for (int i = 1; i <= n; ++i) {
    double x = cos(M_PI * ((double)i - 0.25) / ((double)n + 0.5));
    double v = legendre_eval(n, x);
    double d = legendre_deriv(n, x);
    double dr;
    do {
      dr = v / d;
      x -= dr;
      v = legendre_eval(n, x);
      d = legendre_deriv(n, x);
    } while (fabs(dr) > 2e-16);
    roots[i - 1] = x;
    weights[i - 1] = 2.0 / ((1.0 - x * x) * d * d);
  }
}

/* ChebyshevPolynomial ctor (AngularPolynomials.cpp:99-111). */
void oracle_chebyshev(int n, double *roots, double *angular_roots, double *weights)
{
  for (int i = 1; i <= n; ++i) {
    weights[i - 1] = M_PI / (double)n;
    angular_roots[i - 1] = (2 * (double)i - 1.0) / (2.0 * (double)n) * M_PI;
    roots[i - 1] = cos(angular_roots[i - 1]);
  }
}

/* GCAngularQuadrature::classifyDirection (GCAngularQuadrature.cpp:77-100). */
static int classify(double mu, double eta, double xi)
{
  if (mu > 0.0 && eta > 0.0 && xi > 0.0) return 0;
  if (mu > 0.0 && eta > 0.0 && xi < 0.0) return 1;
  if (mu > 0.0 && eta < 0.0 && xi > 0.0) return 2;
  if (mu > 0.0 && eta < 0.0 && xi < 0.0) return 3;
  if (mu < 0.0 && eta > 0.0 && xi > 0.0) return 4;
  if (mu < 0.0 && eta > 0.0 && xi < 0.0) return 5;
  if (mu < 0.0 && eta < 0.0 && xi > 0.0) return 6;
  if (mu < 0.0 && eta < 0.0 && xi < 0.0) return 7;
  return -1;
}

/* GCAngularQuadrature ctor: product of n_l Legendre polar nodes and n_c
 * Chebyshev azimuths on (0, pi), antipodes added in 3D, then binned by octant
 * keeping insertion order (GCAngularQuadrature.cpp:9-55). */
int oracle_gc_quadrature(int n_c, int n_l, int dims, double *mu, double *eta, double *xi,
                         double *w)
{
  if (n_c <= 0 || n_l <= 0 || (dims != 2 && dims != 3)) return -1;
  double *lr = malloc(sizeof(double) * n_l), *lw = malloc(sizeof(double) * n_l);
  double *cr = malloc(sizeof(double) * n_c), *ca = malloc(sizeof(double) * n_c),
         *cw = malloc(sizeof(double) * n_c);
  oracle_legendre(n_l, lr, lw);
  oracle_chebyshev(n_c, cr, ca, cw);
  int total = n_l * n_c * (dims > 2 ? 2 : 1);
  double *um = malloc(sizeof(double) * total), *ue = malloc(sizeof(double) * total),
         *ux = malloc(sizeof(double) * total), *uw = malloc(sizeof(double) * total);
  int m = 0;
  for (int i = 1; i <= n_l; ++i) {
    for (int j = 1; j <= n_c; ++j) {
      const double weight = lw[i - 1] * cw[j - 1];
      const double u = lr[i - 1];
      const double om = ca[j - 1];
      um[m] = u;
      ue[m] = sqrt(1.0 - (u * u)) * cos(om);
      ux[m] = sqrt(1.0 - (u * u)) * sin(om);
      uw[m] = weight;
      ++m;
      if (dims > 2) {
        um[m] = -1.0 * u;
        ue[m] = -1.0 * sqrt(1.0 - (u * u)) * cos(om);
        ux[m] = -1.0 * sqrt(1.0 - (u * u)) * sin(om);
        uw[m] = weight;
        ++m;
      }
    }
  }
  int per_oct = (n_l * n_c) / 4;
  int counts[8] = {0};
  int ok = 1;
  for (int n = 0; n < m && ok; ++n) {
    int oct = classify(um[n], ue[n], ux[n]);
    if (oct < 0) { ok = 0; break; }
    /* 2D keeps only xi > 0: octants 0,2,4,6 map to slots 0..3. */
    int slot = dims > 2 ? oct : (oct >> 1);
    if (dims == 2 && (oct & 1)) { ok = 0; break; }
    if (counts[slot] >= per_oct) { ok = 0; break; }
    int at = slot * per_oct + counts[slot]++;
    mu[at] = um[n];
    eta[at] = ue[n];
    xi[at] = ux[n];
    w[at] = uw[n];
  }
  int nslots = dims > 2 ? 8 : 4;
  for (int s = 0; s < nslots && ok; ++s)
    if (counts[s] != per_oct) ok = 0;
  free(lr); free(lw); free(cr); free(ca); free(cw);
  free(um); free(ue); free(ux); free(uw);
  return ok ? per_oct : -1;
}

/* GLAngularQuadrature (GLAngularQuadrature.cpp:3-20): the Legendre set itself;
 * TransportSolver1D sweeps [0,n/2) rightwards and [n/2,n) leftwards
 * (TransportSolver1D.cpp:845-860). */
int oracle_gl_quadrature(int n, double *mu, double *w)
{
  if (n <= 0 || (n & 1)) return -1;
  oracle_legendre(n, mu, w);
  return n / 2;
}

/* ------------------------------------------------------------------------- */
/* Cell solves                                                                */
/* ------------------------------------------------------------------------- */

/* TWDiamondDifference3D::solve (TWDiamondDifference3D.h:15-92), theta = theta_s
 * = 0.9, weights from the upwind fluxes when they exceed 1e-8, clamped to
 * [0.5, 1]. */
double oracle_twdd3d_cell(double q, double sigt, double lx, double ly, double lz,
                          double amu, double aeta, double axi, const double in[3],
                          double out[3])
{
  const double theta = 0.9, theta_s = 0.9;
  const double x = in[0], y = in[1], z = in[2];
  const double A = ly * lz;
  const double B = lx * lz;
  const double C = lx * ly;
  const double V = lx * ly * lz;

  double a = 0.5;
  if (x > 1e-8) {
    a = q * V * theta_s + (aeta * B * y + axi * C * z) * theta + amu * A * x;
    a /= (sigt * V + 2.0 * aeta * B + 2.0 * axi * C) * x;
    a = -1.0 * (a - 1.0);
  }
  double b = 0.5;
  if (y > 1e-8) {
    b = q * V * theta_s + (amu * A * x + axi * C * z) * theta + aeta * B * y;
    b /= (sigt * V + 2.0 * amu * A + 2.0 * axi * C) * y;
    b = -1.0 * (b - 1.0);
  }
  double c = 0.5;
  if (z > 1e-8) {
    c = q * V * theta_s + (amu * A * x + aeta * B * y) * theta + axi * C * z;
    c /= (sigt * V + 2.0 * amu * A + 2.0 * aeta * B) * z;
    c = -1.0 * (c - 1.0);
  }
  a = a < 0.5 ? 0.5 : a;
  a = a > 1.0 ? 1.0 : a;
  b = b < 0.5 ? 0.5 : b;
  b = b > 1.0 ? 1.0 : b;
  c = c < 0.5 ? 0.5 : c;
  c = c > 1.0 ? 1.0 : c;

  double psi = q * V + (amu * A / a * x) + (aeta * B / b * y) + (axi * C / c * z);
  psi /= sigt * V + (amu * A / a) + (aeta * B / b) + (axi * C / c);

  out[0] = (psi / a) - (((1.0 - a) / a) * x);
  out[1] = (psi / b) - (((1.0 - b) / b) * y);
  out[2] = (psi / c) - (((1.0 - c) / c) * z);
  return psi;
}

/* DiamondDifference3D::solve (DiamondDifference3D.h:14-49). */
double oracle_dd3d_cell(double q, double sigt, double lx, double ly, double lz,
                        double amu, double aeta, double axi, const double in[3],
                        double out[3])
{
  const double x = in[0], y = in[1], z = in[2];
  double psi = q + (2.0 * x * amu / lx) + (2.0 * y * aeta / ly) + (2.0 * z * axi / lz);
  psi /= (2.0 * amu / lx) + (2.0 * aeta / ly) + (2.0 * axi / lz) + sigt;
  out[0] = 2.0 * psi - x;
  out[1] = 2.0 * psi - y;
  out[2] = 2.0 * psi - z;
  return psi;
}

/* TWDiamondDifference2D::solve (TWDiamondDifference2D.h:15-76), including the
 * "+ abs_xi" term in the b numerator (:51). */
double oracle_twdd2d_cell(double q, double sigt, double lx, double ly, double amu,
                          double aeta, double axi, const double in[2], double out[2])
{
  const double theta = 0.9, theta_s = 0.9;
  const double x = in[0], y = in[1];
  const double A = ly;
  const double B = lx;
  const double V = lx * ly;

  double a = 0.5;
  if (x > 1e-8) {
    a = q * V * theta_s + (aeta * B * y) * theta + amu * A * x;
    a /= (sigt * V + 2.0 * aeta * B) * x;
    a = -1.0 * (a - 1.0);
  }
  double b = 0.5;
  if (y > 1e-8) {
    b = q * V * theta_s + (amu * A * x + axi) * theta + aeta * B * y;
    b /= (sigt * V + 2.0 * amu * A) * y;
    b = -1.0 * (b - 1.0);
  }
  a = a < 0.5 ? 0.5 : a;
  a = a > 1.0 ? 1.0 : a;
  b = b < 0.5 ? 0.5 : b;
  b = b > 1.0 ? 1.0 : b;

  double psi = q * V + (amu * A / a * x) + (aeta * B / b * y);
  psi /= sigt * V + (amu * A / a) + (aeta * B / b);
  out[0] = (psi / a) - (((1.0 - a) / a) * x);
  out[1] = (psi / b) - (((1.0 - b) / b) * y);
  return psi;
}

/* DiamondDifference2D::solve (DiamondDifference2D.h:14-45). */
double oracle_dd2d_cell(double q, double sigt, double lx, double ly, double amu,
                        double aeta, const double in[2], double out[2])
{
  const double x = in[0], y = in[1];
  double psi = q + (2.0 * x * amu / lx) + (2.0 * y * aeta / ly);
  psi /= (2.0 * amu / lx) + (2.0 * aeta / ly) + sigt;
  out[0] = 2.0 * psi - x;
  out[1] = 2.0 * psi - y;
  return psi;
}

/* DiamondDifference1D::solve (DiamondDifference1D.h:14-46); the DSA tallies at
 * :43-45 do not feed back into the flux without DSA and are not restated. */
double oracle_dd1d_cell(double q, double sigt, double lx, double mu, double in, double *out)
{
  const double amu = fabs(mu);
  double psi = q + (2.0 * in * amu / lx);
  psi /= (2.0 * amu / lx) + sigt;
  *out = 2.0 * psi - in;
  return psi;
}

/* ------------------------------------------------------------------------- */
/* Sweeps                                                                     */
/* ------------------------------------------------------------------------- */

typedef struct {
  double *swept;  /* [T][N] per-thread swept scalar flux */
  double *fy;     /* [T][nx] */
  double *fz;     /* [T][nx*ny] */
  double *mom;    /* 1D: [T][N] per-thread current, then [T][N+1] edge scalar fluxes */
} scratch_t;

static long ncells(const oracle_problem *p) { return (long)p->nx * p->ny * p->nz; }

/* sweepPPP..sweepMMM (TransportSolver3D.cpp:655-874): z outer, y, x inner;
 * each loop reversed for a negative direction component.  Octant bits:
 * 4 = -mu, 2 = -eta, 1 = -xi (TransportBase.h:22-32). */
static void sweep3d_dir(const oracle_problem *p, int g, const double *q, int oct, double amu,
                        double aeta, double axi, double wt, double *swept, double *fy,
                        double *fz)
{
  const int nx = p->nx, ny = p->ny, nz = p->nz, G = p->groups;
  const int px = !(oct & 4), py = !(oct & 2), pz = !(oct & 1);
  memset(fz, 0, sizeof(double) * (size_t)nx * ny);
  for (int kk = 0; kk < nz; ++kk) {
    const int k = pz ? kk : nz - 1 - kk;
    memset(fy, 0, sizeof(double) * (size_t)nx);
    for (int jj = 0; jj < ny; ++jj) {
      const int j = py ? jj : ny - 1 - jj;
      double fx = 0.0;
      for (int ii = 0; ii < nx; ++ii) {
        const int i = px ? ii : nx - 1 - ii;
        const long c = ((long)k * ny + j) * nx + i;
        const double sig = p->sigt[(long)p->mat[c] * G + g];
        double in[3] = {fx, fy[i], fz[(long)j * nx + i]}, out[3];
        double psi;
        if (p->scheme == ORACLE_TW_DD)
          psi = oracle_twdd3d_cell(q[c], sig, p->lx[i], p->ly[j], p->lz[k], amu, aeta, axi,
                                   in, out);
        else
          psi = oracle_dd3d_cell(q[c], sig, p->lx[i], p->ly[j], p->lz[k], amu, aeta, axi, in,
                                 out);
        swept[c] += wt * psi;
        fx = out[0];
        fy[i] = out[1];
        fz[(long)j * nx + i] = out[2];
      }
    }
  }
}

/* sweepPPP/PMP/MPP/MMP (TransportSolver2D.cpp:635-725). Slot bits: 2 = -mu,
 * 1 = -eta. */
static void sweep2d_dir(const oracle_problem *p, int g, const double *q, int slot, double amu,
                        double aeta, double axi, double wt, double *swept, double *fy)
{
  const int nx = p->nx, ny = p->ny, G = p->groups;
  const int px = !(slot & 2), py = !(slot & 1);
  memset(fy, 0, sizeof(double) * (size_t)nx);
  for (int jj = 0; jj < ny; ++jj) {
    const int j = py ? jj : ny - 1 - jj;
    double fx = 0.0;
    for (int ii = 0; ii < nx; ++ii) {
      const int i = px ? ii : nx - 1 - ii;
      const long c = (long)j * nx + i;
      const double sig = p->sigt[(long)p->mat[c] * G + g];
      double in[2] = {fx, fy[i]}, out[2];
      double psi;
      if (p->scheme == ORACLE_TW_DD)
        psi = oracle_twdd2d_cell(q[c], sig, p->lx[i], p->ly[j], amu, aeta, axi, in, out);
      else
        psi = oracle_dd2d_cell(q[c], sig, p->lx[i], p->ly[j], amu, aeta, in, out);
      swept[c] += wt * psi;
      fx = out[0];
      fy[i] = out[1];
    }
  }
}

/* sweepR / sweepL (TransportSolver1D.cpp:890-915).  With cur (DSA), the tallies
 * of DiamondDifference1D.h:41-43: w mu psi into the cell current and w psi (the
 * cell-centre flux) onto the downwind edge, edge i+1 for mu > 0 and edge i for
 * mu < 0 (CartesianCell1D.cpp:116-120). */
static void sweep1d_dir(const oracle_problem *p, int g, const double *q, int half, double mu,
                        double wt, double *swept, double *cur, double *edge)
{
  const int nx = p->nx, G = p->groups;
  double f = 0.0;
  for (int ii = 0; ii < nx; ++ii) {
    const int i = half == 0 ? ii : nx - 1 - ii;
    const double sig = p->sigt[(long)p->mat[i] * G + g];
    double out;
    const double psi = oracle_dd1d_cell(q[i], sig, p->lx[i], mu, f, &out);
    swept[i] += wt * psi;
    if (cur) {
      edge[half == 0 ? i + 1 : i] += wt * psi;
      cur[i] += wt * mu * psi;
    }
    f = out;
  }
}

static int team(int nthreads) { return nthreads < 1 ? 1 : nthreads; }

static int sweep_with(const oracle_problem *p, int g, const double *q, double *dphi, int T,
                      scratch_t *s, double *mom_out)
{
  const long N = ncells(p);
  const int D = p->dirs;
  if (p->dims == 3) {
    /* TransportSolver3D::sweep / parallelSweepOctant (:565-653). */
    for (int oct = 0; oct < 8; ++oct) {
#pragma omp parallel for schedule(static) num_threads(T)
      for (int n = 0; n < D; ++n) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        const int at = oct * D + n;
        sweep3d_dir(p, g, q, oct, fabs(p->mu[at]), fabs(p->eta[at]), fabs(p->xi[at]), p->w[at],
                    s->swept + (size_t)tid * N, s->fy + (size_t)tid * p->nx,
                    s->fz + (size_t)tid * p->nx * p->ny);
      }
    }
  } else if (p->dims == 2) {
    for (int slot = 0; slot < 4; ++slot) {
#pragma omp parallel for schedule(static) num_threads(T)
      for (int n = 0; n < D; ++n) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        const int at = slot * D + n;
        sweep2d_dir(p, g, q, slot, fabs(p->mu[at]), fabs(p->eta[at]), fabs(p->xi[at]), p->w[at],
                    s->swept + (size_t)tid * N, s->fy + (size_t)tid * p->nx);
      }
    }
  } else if (p->dims == 1) {
    for (int half = 0; half < 2; ++half) {
#pragma omp parallel for schedule(static) num_threads(T)
      for (int n = 0; n < D; ++n) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        const int at = half * D + n;
        double *cur = mom_out ? s->mom + (size_t)tid * N : NULL;
        double *edge = mom_out ? s->mom + (size_t)T * N + (size_t)tid * (N + 1) : NULL;
        sweep1d_dir(p, g, q, half, p->mu[at], p->w[at], s->swept + (size_t)tid * N, cur, edge);
      }
    }
  } else {
    return -1;
  }
  /* Thread-ordered reduction of the swept partials (TransportSolver3D.cpp:609-621). */
#pragma omp parallel for schedule(static) num_threads(T)
  for (long c = 0; c < N; ++c) {
    double acc = 0.0;
    for (int t = 0; t < T; ++t) {
      acc += s->swept[(size_t)t * N + c];
      s->swept[(size_t)t * N + c] = 0.0;
    }
    dphi[c] = acc;
  }
  if (mom_out) {
    /* moments (TransportSolver1D.cpp:862-882): [0] cell current, [1] scalar flux
     * of the cell's right edge, [2] of its left edge, each summed over threads */
    const double *edge = s->mom + (size_t)T * N;
    for (long c = 0; c < N; ++c) {
      double j = 0.0, r = 0.0, l = 0.0;
      for (int t = 0; t < T; ++t) {
        j += s->mom[(size_t)t * N + c];
        r += edge[(size_t)t * (N + 1) + c + 1];
        l += edge[(size_t)t * (N + 1) + c];
      }
      mom_out[c] = j;
      mom_out[N + c] = r;
      mom_out[2 * N + c] = l;
    }
    memset(s->mom, 0, sizeof(double) * (size_t)T * (2 * N + 1));
  }
  return 0;
}

static int scratch_alloc(const oracle_problem *p, int T, scratch_t *s)
{
  const long N = ncells(p);
  s->swept = calloc((size_t)T * N, sizeof(double));
  s->fy = calloc((size_t)T * p->nx, sizeof(double));
  s->fz = calloc((size_t)T * p->nx * p->ny, sizeof(double));
  s->mom = p->dims == 1 ? calloc((size_t)T * (2 * N + 1), sizeof(double)) : NULL;
  return (s->swept && s->fy && s->fz && (p->dims != 1 || s->mom)) ? 0 : -1;
}

static void scratch_free(scratch_t *s)
{
  free(s->swept);
  free(s->fy);
  free(s->fz);
  free(s->mom);
}

int oracle_sweep(const oracle_problem *p, int g, const double *q, double *dphi, int nthreads)
{
  const int T = team(nthreads);
  scratch_t s;
  if (scratch_alloc(p, T, &s)) { scratch_free(&s); return -1; }
  int rc = sweep_with(p, g, q, dphi, T, &s, NULL);
  scratch_free(&s);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* Source iteration and multigroup Gauss-Seidel                               */
/* ------------------------------------------------------------------------- */

/* Per-cell measure used by the residuals: volume (3D), area (2D), length (1D). */
static double measure(const oracle_problem *p, long c)
{
  const int i = (int)(c % p->nx);
  const int j = (int)((c / p->nx) % p->ny);
  const int k = (int)(c / ((long)p->nx * p->ny));
  if (p->dims == 3) return p->lx[i] * p->ly[j] * p->lz[k];
  if (p->dims == 2) return p->lx[i] * p->ly[j];
  return p->lx[i];
}

/* Angular normalisation of isotropic sources: 0.25/pi (3D,
 * TransportSolver3D.cpp:351,375,542), 0.5/pi (2D, TransportSolver2D.cpp:351,375,
 * 542), 0.5 (1D, TransportSolver1D.cpp:384,408,814). */
static double norm_term(const oracle_problem *p, double a, double b)
{
  if (p->dims == 3) return 0.25 * a * b / M_PI;
  if (p->dims == 2) return 0.5 * a * b / M_PI;
  return 0.5 * a * b;
}
static double norm_src(const oracle_problem *p, double a)
{
  if (p->dims == 3) return 0.25 * a / M_PI;
  if (p->dims == 2) return 0.5 * a / M_PI;
  return 0.5 * a;
}

typedef struct {
  const oracle_problem *p;
  int T;
  scratch_t s;
  double *q;     /* current iteration source */
  double *dphi;  /* current scalar flux (increment in 2D/3D) */
  double *qmg;   /* 1D multigroup source (_mg_source) */
  double *phi;   /* [G][N] total scalar flux */
  double *prev;  /* [G][N] previous MG iterate */
  double *mom;   /* 1D DSA: [3][N] current, right and left interface fluxes */
  int eigen;     /* solving the eigenvalue problem (selects the DSA scheme) */
  const oracle_problem *ps;  /* problem the sweeps see (sigma_t + inv_v/dt in a transient) */
  int transient;
  double t, dt;
  double *last;  /* [G][N] previous step's flux */
  double *dnp;   /* [8][N] precursors */
  int sweeps;
} solver_t;

/* Prompt fission term: 0.25 (2D, 1D: 0.5) * f * chi_g * nuSf_g' * phi_g', / pi in
 * 2D/3D, with f = 1 - sum_d beta (TransportSolver3D.cpp:377-387,
 * TransportSolver2D.cpp:377-387, TransportSolver1D.cpp:410-420). */
static double fission_term(const oracle_problem *p, double f, double chi, double prod,
                           double phi)
{
  const double t = (p->dims == 3 ? 0.25 : 0.5) * f * chi * prod * phi;
  return p->dims == 1 ? t : t / M_PI;
}

/* updateMultigroupSource (TransportSolver3D.cpp:339-402, TransportSolver2D.cpp:
 * 339-402, TransportSolver1D.cpp:369-435): external source, in-scatter and prompt
 * fission.  With eigen != 0, updateMultigroupSourceEigen (TransportSolver3D.cpp:
 * 404-430, TransportSolver2D.cpp:404-430, TransportSolver1D.cpp:437-466):
 * in-scatter and fission / k on top of the iteration source left by the last
 * scattering update in 2D/3D (which the reference does not clear), from zero in 1D. */
static void update_mg_source_k(solver_t *S, int g, int eigen, double k)
{
  const oracle_problem *p = S->p;
  const long N = ncells(p);
  const int G = p->groups;
  for (long c = 0; c < N; ++c) {
    const int m = p->mat[c];
    const int fis = p->chi && p->has_chi[m];
    double acc;
    if (eigen)
      acc = p->dims == 1 ? 0.0 : S->q[c];
    else
      acc = p->has_src[m] ? norm_src(p, p->src[(long)m * G + g]) : 0.0;
    const double nrm = p->dims == 3 ? 0.25 : 0.5;
    /* step source at the step's time (TransportSolver3D.cpp:353-369) */
    if (S->transient && p->step_type && p->step_type[m] >= 0) {
      const int ty = p->step_type[m];
      const double a = p->step_in[m], b = p->step_out[m], t = S->t;
      const int on = ty == 0 ? (a <= t && t <= b) : (ty == 1 ? a <= t : t <= b);
      acc += on ? norm_src(p, p->step_src[(long)m * G + g]) : 0.0;
    }
    const int nd = S->transient ? p->nd[m] : 0;
    for (int gp = 0; gp < G; ++gp) {
      const double f = S->phi[(long)gp * N + c];
      if (gp != g)
        acc += norm_term(p, p->sigs[((long)m * G + g) * G + gp], f);
      if (fis) {
        const double chi = p->chi[(long)m * G + g], pr = p->prod[(long)m * G + gp];
        if (eigen)
          acc += fission_term(p, 1.0, chi, pr, f) / k;
        else
          acc += fission_term(p, p->prompt[(long)m * G + gp], chi, pr, f);
      }
      /* delayed and time-derivative sources, inside the g' loop as the reference
       * has them (TransportSolver3D.cpp:388-396) */
      for (int d = 0; d < nd; ++d) {
        const double v = nrm * p->chi_d[((long)m * G + g) * ORACLE_MAX_DELAYED + d] *
                         S->dnp[(long)d * N + c] * p->lambda[(long)m * ORACLE_MAX_DELAYED + d];
        acc += p->dims == 1 ? v : v / M_PI;
      }
      if (S->transient) {
        const double v = nrm * S->last[(long)g * N + c] * p->inv_v[(long)m * G + g] / S->dt;
        acc += p->dims == 1 ? v : v / M_PI;
      }
    }
    S->dphi[c] = 0.0;
    if (p->dims == 1)
      S->qmg[c] = acc;
    else
      S->q[c] = acc;
    S->prev[(long)g * N + c] = S->phi[(long)g * N + c];
    S->phi[(long)g * N + c] = 0.0;
  }
}

static void update_mg_source(solver_t *S, int g) { update_mg_source_k(S, g, 0, 0.0); }

/* computeScatteringResidual: increment form in 2D/3D (TransportSolver3D.cpp:
 * 547-563), standard form in 1D (TransportSolver1D.cpp:822-838). */
static double scattering_residual(solver_t *S, int g)
{
  const oracle_problem *p = S->p;
  const long N = ncells(p);
  double diff = 0.0, total = 0.0;
  for (long c = 0; c < N; ++c) {
    const double v = measure(p, c);
    const double d = S->dphi[c];
    const double t = S->phi[(long)g * N + c];
    if (p->dims == 1) {
      diff = diff + (d - t) * (d - t) * v;
      total = total + d * d * v;
    } else {
      diff = diff + d * d * v;
      total = total + (t + d) * (t + d) * v;
    }
  }
  return total > 1e-8 ? sqrt(diff) / sqrt(total) : 0.0;
}

/* updateScatteringSource (TransportSolver3D.cpp:531-545 increment form;
 * TransportSolver1D.cpp:802-820 standard form). */
static void update_scattering_source(solver_t *S, int g)
{
  const oracle_problem *p = S->p;
  const long N = ncells(p);
  const int G = p->groups;
  for (long c = 0; c < N; ++c) {
    const double sgg = p->sigs[((long)p->mat[c] * G + g) * G + g];
    if (p->dims == 1) {
      S->phi[(long)g * N + c] = S->dphi[c];
      S->q[c] = S->qmg[c];
      S->q[c] += 0.5 * sgg * S->dphi[c];
    } else {
      S->phi[(long)g * N + c] += S->dphi[c];
      S->q[c] = norm_term(p, sgg, S->dphi[c]);
    }
    S->dphi[c] = 0.0;
  }
}

void oracle_tridiag_solve(int n, double *lo, double *di, double *up, double *rhs, double *x)
{
  double *up2 = calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (int i = 0; i < n - 1; ++i) {
    if (fabs(di[i]) >= fabs(lo[i + 1])) {
      const double f = lo[i + 1] / di[i];
      di[i + 1] -= f * up[i];
      rhs[i + 1] -= f * rhs[i];
    } else {
      /* interchange rows i and i+1, then eliminate below */
      const double f = di[i] / lo[i + 1];
      const double d1 = di[i + 1], u1 = i < n - 2 ? up[i + 1] : 0.0, r0 = rhs[i];
      di[i] = lo[i + 1];
      di[i + 1] = up[i] - f * d1;
      up[i] = d1;
      if (i < n - 2) {
        up2[i] = u1;
        up[i + 1] = -f * u1;
      }
      rhs[i] = rhs[i + 1];
      rhs[i + 1] = r0 - f * rhs[i];
    }
  }
  x[n - 1] = rhs[n - 1] / di[n - 1];
  if (n > 1) x[n - 2] = (rhs[n - 2] - up[n - 2] * x[n - 1]) / di[n - 2];
  for (int i = n - 3; i >= 0; --i)
    x[i] = (rhs[i] - up[i] * x[i + 1] - up2[i] * x[i + 2]) / di[i];
  free(up2);
}

/* The diffusion system of syntheticAcceleration (TransportSolver1D.cpp:587-800):
 * Alcouffe's equations on the N+1 cell edges from a sweep's flux phi, cell
 * currents cur and edge fluxes sfr/sfl (right/left edge of each cell); fixed-
 * source mode: source correction (Eq. 22-23); eigenvalue mode: modified
 * diffusion coefficients D = -h J / (sfR - sfL), or, if any is <= 0, the removal
 * correction.  Boundary rows after Larsen (Eq. 34, 37). */
void oracle_dsa_system(const oracle_problem *p, int g, int eigen, const double *phi,
                       const double *cur, const double *sfr, const double *sfl,
                       const double *qmg, double *lo, double *di, double *up, double *rhs)
{
  const int N = p->nx, G = p->groups, D = p->dirs;
  const double *l = p->lx;
  memset(lo, 0, sizeof(double) * (size_t)(N + 1));
  memset(up, 0, sizeof(double) * (size_t)(N + 1));
#define TOT(c) (p->sigt[(long)p->mat[c] * G + g])
#define SELF(c) (p->sigs[((long)p->mat[c] * G + g) * G + g])
  int removal = 0;
  if (eigen) {
    for (int i = 0; i < N - 1; ++i) {
      double Di = -1.0 * l[i] * cur[i];
      Di /= (sfr[i] - sfl[i]);
      double Di1 = -1.0 * l[i + 1] * cur[i + 1];
      Di1 /= (sfr[i + 1] - sfl[i + 1]);
      if (Di <= 0.0 || Di1 <= 0.0) { removal = 1; break; }
      double sr = (TOT(i) - SELF(i)) * l[i] * phi[i];
      sr += (TOT(i + 1) - SELF(i + 1)) * l[i + 1] * phi[i + 1];
      sr *= 0.5 / sfr[i];
      lo[i + 1] = -1.0 * Di / l[i];
      di[i + 1] = (Di / l[i]) + (Di1 / l[i + 1]) + sr;
      up[i + 1] = -1.0 * Di1 / l[i + 1];
      rhs[i + 1] = (qmg[i] * l[i] + qmg[i + 1] * l[i + 1]);
    }
  }
  if (!eigen || removal) {
    for (int i = 0; i < N - 1; ++i) {
      const double Dh = 1.0 / 3.0 / TOT(i) / l[i];
      const double Dh1 = 1.0 / 3.0 / TOT(i + 1) / l[i + 1];
      double sr = (TOT(i) - SELF(i)) * l[i] * phi[i];
      sr += (TOT(i + 1) - SELF(i + 1)) * l[i + 1] * phi[i + 1];
      sr *= 0.5 / sfr[i];
      double r = cur[i + 1] - cur[i];
      r += Dh1 * (sfr[i + 1] - sfl[i + 1]);
      r -= Dh * (sfr[i] - sfl[i]);
      const double qqh = (qmg[i] * l[i] + qmg[i + 1] * l[i + 1]);
      lo[i + 1] = -1.0 * Dh;
      up[i + 1] = -1.0 * Dh1;
      if (eigen) {
        r /= sfr[i];
        di[i + 1] = Dh + Dh1 + sr + r;
        rhs[i + 1] = qqh;
      } else {
        di[i + 1] = Dh + Dh1 + sr;
        rhs[i + 1] = qqh - r;
      }
    }
  }
  double bl = 0.0, br = 0.0;
  for (int t = 0; t < D; ++t) bl += p->mu[t] * p->w[t];
  for (int t = D; t < 2 * D; ++t) br += p->mu[t] * p->w[t];
  {
    rhs[0] = -1.0 * qmg[0];
    const double hd = 0.5 / 3.0 / TOT(0) / l[0];
    const double es = 0.125 * (TOT(0) - SELF(0)) * l[0];
    di[0] = bl + hd + es;
    up[0] = -1.0 * hd + es;
  }
  {
    const int I = N - 1;
    rhs[N] = qmg[I];
    const double hd = 0.5 / 3.0 / TOT(I) / l[I];
    const double es = 0.125 * (TOT(I) - SELF(I)) * l[I];
    di[N] = br - hd - es;
    lo[N] = hd - es;
  }
#undef TOT
#undef SELF
}

/* syntheticAcceleration: solve the system; the cell flux becomes the average of
 * its two edge values (TransportSolver1D.cpp:797-799). */
static void synthetic_acceleration(solver_t *S, int g)
{
  const oracle_problem *p = S->p;
  const int N = p->nx, n = N + 1;
  double *lo = calloc((size_t)n, sizeof(double)), *di = calloc((size_t)n, sizeof(double));
  double *up = calloc((size_t)n, sizeof(double)), *rhs = calloc((size_t)n, sizeof(double));
  double *x = calloc((size_t)n, sizeof(double));
  oracle_dsa_system(p, g, S->eigen, S->dphi, S->mom, S->mom + N, S->mom + 2 * N, S->qmg, lo, di,
                    up, rhs);
  oracle_tridiag_solve(n, lo, di, up, rhs, x);
  for (int c = 0; c < N; ++c) S->dphi[c] = 0.5 * (x[c] + x[c + 1]);
  free(lo); free(di); free(up); free(rhs); free(x);
}

/* sourceIteration (TransportSolver3D.cpp:458-497; TransportSolver1D.cpp:495-541,
 * which first applies updateScatteringSource once). */
static int source_iteration(solver_t *S, int g, double tol, int max_it, int *its, double *res,
                            double *trace)
{
  const oracle_problem *p = S->p;
  if (p->dims == 1) update_scattering_source(S, g);
  int it = 0;
  double r = 0.0;
  do {
    const int dsa = p->dims == 1 && p->dsa;
    sweep_with(S->ps, g, S->q, S->dphi, S->T, &S->s, dsa ? S->mom : NULL);
    if (dsa) synthetic_acceleration(S, g);
    S->sweeps++;
    r = scattering_residual(S, g);
    if (trace) trace[it] = r;
    update_scattering_source(S, g);
    it++;
  } while (it < max_it && tol < r);
  if (its) *its = it;
  if (res) *res = r;
  return it < max_it;
}

/* computeMGFluxResidual (TransportSolver3D.cpp:318-337). */
static double mg_residual(solver_t *S)
{
  const oracle_problem *p = S->p;
  const long N = ncells(p);
  const int G = p->groups;
  double diff = 0.0, total = 0.0;
  for (long c = 0; c < N; ++c) {
    const double v = measure(p, c);
    for (int g = 0; g < G; ++g) {
      const double f = S->phi[(long)g * N + c];
      const double d = f - S->prev[(long)g * N + c];
      diff = diff + d * d * v;
      total = total + f * f * v;
    }
  }
  return total > 1e-8 ? sqrt(diff) / sqrt(total) : 0.0;
}

static int solver_init(solver_t *S, const oracle_problem *p, int nthreads)
{
  const long N = ncells(p);
  memset(S, 0, sizeof(*S));
  S->p = p;
  S->ps = p;
  S->T = team(nthreads);
  S->q = calloc(N, sizeof(double));
  S->dphi = calloc(N, sizeof(double));
  S->qmg = calloc(N, sizeof(double));
  S->phi = calloc((size_t)p->groups * N, sizeof(double));
  S->prev = calloc((size_t)p->groups * N, sizeof(double));
  S->mom = calloc((size_t)3 * N, sizeof(double));
  if (!S->q || !S->dphi || !S->qmg || !S->phi || !S->prev || !S->mom) return -1;
  return scratch_alloc(p, S->T, &S->s);
}

static void solver_free(solver_t *S)
{
  free(S->q); free(S->dphi); free(S->qmg); free(S->phi); free(S->prev); free(S->mom);
  scratch_free(&S->s);
}

/* solveFixedSource (TransportSolver3D.cpp:262-316): Gauss-Seidel over groups,
 * each group restarted from zero, stop when G == 1, the MG residual <= mg_tol or
 * mg_max_it passes are done; success iff fewer than mg_max_it passes. */
int oracle_solve_fixed_source(const oracle_problem *p, double fs_tol, int fs_max_it,
                              double mg_tol, int mg_max_it, int nthreads, double *phi_out,
                              oracle_stats *stats)
{
  solver_t S;
  if (solver_init(&S, p, nthreads)) { solver_free(&S); return -1; }
  const long N = ncells(p);
  const int G = p->groups;
  int mg_it = 0, ok = 1;
  double r = 0.0;
  if (stats) memset(stats, 0, sizeof(*stats));
  do {
    for (int g = 0; g < G; ++g) {
      update_mg_source(&S, g);
      int its = 0;
      double res = 0.0;
      int conv = source_iteration(&S, g, fs_tol, fs_max_it, &its, &res, NULL);
      if (stats && g < 64) {
        stats->si_iterations_last[g] = its;
        stats->si_residual_last[g] = res;
      }
      if (!conv) { ok = 0; break; }
    }
    if (!ok) break;
    r = mg_residual(&S);
    mg_it++;
  } while (mg_it < mg_max_it && mg_tol < r && G > 1);
  int success = ok && (mg_it < mg_max_it);
  if (phi_out) memcpy(phi_out, S.phi, sizeof(double) * (size_t)G * N);
  if (stats) {
    stats->success = success;
    stats->mg_iterations = mg_it;
    stats->mg_residual = r;
    stats->sweeps = S.sweeps;
  }
  solver_free(&S);
  return success;
}

/* updateEigenvalue (TransportSolver3D.cpp:432-456, TransportSolver2D.cpp:432-456,
 * TransportSolver1D.cpp:468-493): returns the new k. */
static double update_eigenvalue(solver_t *S, double k)
{
  const oracle_problem *p = S->p;
  const long N = ncells(p);
  const int G = p->groups;
  double num = 0.0, den = 0.0;
  for (long c = 0; c < N; ++c) {
    const int m = p->mat[c];
    if (!p->has_prod[m]) continue;
    const double v = measure(p, c);
    for (int g = 0; g < G; ++g) {
      const double pr = p->prod[(long)m * G + g];
      num = num + v * pr * S->phi[(long)g * N + c];
      if (p->dims == 1)
        den = den + v * pr * S->prev[(long)g * N + c];
      else
        den = den + v * pr * S->prev[(long)g * N + c] / k;
    }
  }
  return p->dims == 1 ? k * num / den : num / den;
}

/* solveEigenvalue (TransportSolver3D.cpp:29-86; TransportSolver1D.cpp:36-90):
 * phi = 1, k = 1, k_prev = 0 (initializeSolve, constructor); per pass the eigen
 * source and source iteration of every group, then k and the shape residual;
 * a failed source iteration ends the solve with false. */
int oracle_solve_eigenvalue(const oracle_problem *p, double fs_tol, int fs_max_it,
                            double mg_tol, int mg_max_it, double k_tol, int nthreads,
                            double *phi_out, double *k_eff, oracle_stats *stats)
{
  if (!p->chi || !p->prod || !p->has_chi || !p->has_prod) return -1;
  solver_t S;
  if (solver_init(&S, p, nthreads)) { solver_free(&S); return -1; }
  const long N = ncells(p);
  const int G = p->groups;
  for (long e = 0; e < (long)G * N; ++e) S.phi[e] = 1.0;
  S.eigen = 1;
  double k = 1.0, k_prev = 0.0, shape = 0.0, kdiff = 0.0;
  int mg_it = 0, ok = 1;
  if (stats) memset(stats, 0, sizeof(*stats));
  do {
    for (int g = 0; g < G; ++g) {
      update_mg_source_k(&S, g, 1, k);
      int its = 0;
      double res = 0.0;
      int conv = source_iteration(&S, g, fs_tol, fs_max_it, &its, &res, NULL);
      if (stats && g < 64) {
        stats->si_iterations_last[g] = its;
        stats->si_residual_last[g] = res;
      }
      if (!conv) { ok = 0; break; }
    }
    if (!ok) break;
    k_prev = k;
    k = update_eigenvalue(&S, k);
    shape = mg_residual(&S);
    kdiff = fabs(k - k_prev);
    mg_it++;
  } while (mg_it < mg_max_it && (mg_tol < shape || k_tol < kdiff));
  int success = ok && (mg_it < mg_max_it);
  if (phi_out) memcpy(phi_out, S.phi, sizeof(double) * (size_t)G * N);
  if (k_eff) *k_eff = k;
  if (stats) {
    stats->success = success;
    stats->mg_iterations = mg_it;
    stats->mg_residual = shape;
    stats->sweeps = S.sweeps;
    stats->k_eff = k;
    stats->k_residual = kdiff;
  }
  solver_free(&S);
  return success;
}

/* stepDNPs (TransportSolver3D.cpp:230-260). */
static void step_dnps(solver_t *S)
{
  const oracle_problem *p = S->p;
  const long N = ncells(p);
  const int G = p->groups, K = ORACLE_MAX_DELAYED;
  for (long c = 0; c < N; ++c) {
    const int m = p->mat[c];
    for (int d = 0; d < p->nd[m]; ++d) {
      double src = 0.0;
      for (int g = 0; g < G; ++g)
        src += p->prod[(long)m * G + g] * S->phi[(long)g * N + c] *
               p->beta[((long)m * G + g) * K + d];
      const double lam = p->lambda[(long)m * K + d];
      const double e = exp(-lam * S->dt);
      S->dnp[(long)d * N + c] = (src / lam) * (1.0 - e) + S->dnp[(long)d * N + c] * e;
    }
  }
}

int oracle_solve_transient(const oracle_problem *p, double fs_tol, int fs_max_it, double mg_tol,
                           int mg_max_it, double t0, double t1, int steps, int nthreads,
                           double *phi_steps, double *dnp_steps, oracle_stats *stats)
{
  if (!p->inv_v || !p->nd || steps < 1) return -1;
  solver_t S;
  if (solver_init(&S, p, nthreads)) { solver_free(&S); return -1; }
  const long N = ncells(p);
  const int G = p->groups, K = ORACLE_MAX_DELAYED;
  /* the sweeps see sigma_t + inv_v / dt (DiamondDifference3D.h:25-26) */
  S.dt = (t1 - t0) / (double)steps;
  double *sig = malloc(sizeof(double) * (size_t)p->num_mats * G);
  for (long e = 0; e < (long)p->num_mats * G; ++e) sig[e] = p->sigt[e] + p->inv_v[e] / S.dt;
  oracle_problem ps = *p;
  ps.sigt = sig;
  S.ps = &ps;
  S.transient = 1;
  S.last = calloc((size_t)G * N, sizeof(double));
  S.dnp = calloc((size_t)K * N, sizeof(double));
  S.t = 0.0;
  if (stats) memset(stats, 0, sizeof(*stats));
  int all_ok = 1;
  for (int step = 0; step < steps; ++step) {
    int mg_it = 0, ok = 1;
    double r = 0.0;
    do {
      for (int g = 0; g < G; ++g) {
        update_mg_source(&S, g);
        int its = 0;
        double res = 0.0;
        int conv = source_iteration(&S, g, fs_tol, fs_max_it, &its, &res, NULL);
        if (stats && g < 64) {
          stats->si_iterations_last[g] = its;
          stats->si_residual_last[g] = res;
        }
        if (!conv) { ok = 0; break; }
      }
      if (!ok) break;
      r = mg_residual(&S);
      mg_it++;
    } while (mg_it < mg_max_it && mg_tol < r && G > 1);
    if (stats) {
      stats->success = ok && mg_it < mg_max_it;
      stats->mg_iterations = mg_it;
      stats->mg_residual = r;
      stats->sweeps = S.sweeps;
    }
    if (!(ok && mg_it < mg_max_it)) { all_ok = 0; break; }
    step_dnps(&S);
    if (phi_steps) memcpy(phi_steps + (size_t)step * G * N, S.phi, sizeof(double) * (size_t)G * N);
    if (dnp_steps) memcpy(dnp_steps + (size_t)step * K * N, S.dnp, sizeof(double) * (size_t)K * N);
    /* updateStepFluxes (TransportSolver3D.cpp:204-228) */
    memcpy(S.last, S.phi, sizeof(double) * (size_t)G * N);
    memset(S.phi, 0, sizeof(double) * (size_t)G * N);
    S.t += S.dt;
  }
  free(sig); free(S.last); free(S.dnp);
  solver_free(&S);
  return all_ok;
}

int oracle_source_iteration(const oracle_problem *p, int g, const double *qmg, double fs_tol,
                            int fs_max_it, int nthreads, double *phi_g, int *iterations,
                            double *residual, double *residual_trace)
{
  solver_t S;
  if (solver_init(&S, p, nthreads)) { solver_free(&S); return -1; }
  const long N = ncells(p);
  if (p->dims == 1)
    memcpy(S.qmg, qmg, sizeof(double) * N);
  else
    memcpy(S.q, qmg, sizeof(double) * N);
  int conv = source_iteration(&S, g, fs_tol, fs_max_it, iterations, residual, residual_trace);
  if (phi_g) memcpy(phi_g, S.phi + (long)g * N, sizeof(double) * N);
  solver_free(&S);
  return conv;
}
