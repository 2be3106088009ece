// iterate.hip -- source iteration and multigroup bookkeeping kernels.
//
// All reductions are deterministic: every block walks a fixed set of cells in a
// fixed order and the block partials are summed by one block in index order, so
// residuals (and therefore iteration counts) are reproducible run to run.
// Arithmetic keeps the reference's expression order (no contraction) so that the
// FP64 bookkeeping matches the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.h"

namespace doma
{
namespace
{
constexpr int RED_THREADS = 256;
constexpr double PI = 3.14159265358979323846;

// Angular normalisation of isotropic sources: 3D 0.25/pi (TransportSolver3D.cpp:
// 351, 375, 542), 2D 0.5/pi (TransportSolver2D.cpp:351, 375, 542), 1D 0.5
// (TransportSolver1D.cpp:384, 408, 814).
__device__ __forceinline__ double
norm_src(int dims, double a)
{
#pragma clang fp contract(off)
  if (dims == 3)
    return 0.25 * a / PI;
  if (dims == 2)
    return 0.5 * a / PI;
  return 0.5 * a;
}

__device__ __forceinline__ double
norm_term(int dims, double a, double b)
{
#pragma clang fp contract(off)
  if (dims == 3)
    return 0.25 * a * b / PI;
  if (dims == 2)
    return 0.5 * a * b / PI;
  return 0.5 * a * b;
}

__device__ __forceinline__ double
measure(const IterArgs & a, int c)
{
#pragma clang fp contract(off)
  const int i = c % a.nx;
  const int j = (c / a.nx) % a.ny;
  const int k = c / (a.nx * a.ny);
  if (a.dims == 3)
    return a.lx[i] * a.ly[j] * a.lz[k];
  if (a.dims == 2)
    return a.lx[i] * a.ly[j];
  return a.lx[i];
}

// Fixed-order block reduction of two sums; thread 0 gets the totals.
__device__ __forceinline__ void
block_sum2(double & u, double & v)
{
  __shared__ double su[RED_THREADS], sv[RED_THREADS];
  su[threadIdx.x] = u;
  sv[threadIdx.x] = v;
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1)
  {
    if ((int)threadIdx.x < s)
    {
      su[threadIdx.x] += su[threadIdx.x + s];
      sv[threadIdx.x] += sv[threadIdx.x + s];
    }
    __syncthreads();
  }
  u = su[0];
  v = sv[0];
}

// Prompt fission term of group g from group g' (TransportSolver3D.cpp:377-387,
// TransportSolver2D.cpp:377-387, TransportSolver1D.cpp:410-420):
// norm * f * chi_g * nuSf_g' * phi_g' (/ pi in 2D/3D), f the prompt fraction
// (1 - sum_d beta, exactly 1 without delayed groups).
__device__ __forceinline__ double
fission_term(int dims, double f, double chi, double prod, double phi)
{
#pragma clang fp contract(off)
  double t = (dims == 3 ? 0.25 : 0.5) * f;
  t = t * chi;
  t = t * prod;
  t = t * phi;
  return dims == 1 ? t : t / PI;
}

// updateMultigroupSource (TransportSolver3D.cpp:339-402; 2D :339-402; 1D
// TransportSolver1D.cpp:369-435), steady-state terms (external source,
// in-scatter, prompt fission), fused with the gather of the group's total cross
// section.  With a.eigen, updateMultigroupSourceEigen (TransportSolver3D.cpp:
// 404-430; 2D :404-430; 1D TransportSolver1D.cpp:437-466): in-scatter plus
// fission / k and no external source.  In 2D/3D the reference does not clear the
// iteration source first, so the previous group's last scattering update (left
// in q by the source iteration) stays in it; 1D starts from zero.
__global__ void
mg_source_kernel(IterArgs a)
{
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  const int m = a.mat[c];
  const int G = a.G, g = a.g;
  const bool fis = a.chi && a.has_chi[m];
  const bool tr = a.inv_v != nullptr;
  double acc;
  if (a.eigen)
    acc = a.dims == 1 ? 0.0 : a.q[c];
  else
    acc = a.has_src[m] ? norm_src(a.dims, a.src[(size_t)m * G + g]) : 0.0;
  // Step source of the cell's block at the step's time (TransportSolver3D.cpp:353-369).
  if (tr && a.step_type && a.step_type[m] != DOMA_STEP_NONE)
  {
    const int ty = a.step_type[m];
    const double tin = a.step_in[m], tout = a.step_out[m];
    const bool on = ty == DOMA_STEP_BOTH ? (tin <= a.t && a.t <= tout)
                                         : (ty == DOMA_STEP_INSERT ? tin <= a.t : a.t <= tout);
    acc += on ? norm_src(a.dims, a.step_src[(size_t)m * G + g]) : 0.0;
  }
  const double chi = fis ? a.chi[(size_t)m * G + g] : 0.0;
  const int nd = tr ? a.nd[m] : 0;
  for (int gp = 0; gp < G; ++gp)
  {
    const double f = a.phi[(size_t)gp * a.N + c];
    if (gp != g)
      acc += norm_term(a.dims, a.sigs[((size_t)m * G + g) * G + gp], f);
    if (fis)
    {
      const double pr = a.prod[(size_t)m * G + gp];
      if (a.eigen)
        acc += fission_term(a.dims, 1.0, chi, pr, f) / a.k;
      else
        acc += fission_term(a.dims, a.prompt[(size_t)m * G + gp], chi, pr, f);
    }
    // The reference adds the delayed and the time-derivative sources inside the
    // loop over source groups, i.e. G times each (TransportSolver3D.cpp:388-396).
    for (int d = 0; d < nd; ++d)
    {
      double t = (a.dims == 3 ? 0.25 : 0.5) * a.chi_d[((size_t)m * G + g) * DOMA_MAX_DELAYED + d];
      t = t * a.dnp[(size_t)d * a.N + c];
      t = t * a.lambda[(size_t)m * DOMA_MAX_DELAYED + d];
      acc += a.dims == 1 ? t : t / PI;
    }
    if (tr)
    {
      double t = (a.dims == 3 ? 0.25 : 0.5) * a.last[(size_t)g * a.N + c];
      t = t * a.inv_v[(size_t)m * G + g];
      t = t / a.dt;
      acc += a.dims == 1 ? t : t / PI;
    }
  }
  if (a.dims == 1)
    a.qmg[c] = acc;
  else
    a.q[c] = acc;
  a.prev[(size_t)g * a.N + c] = a.phi[(size_t)g * a.N + c];
  a.phi[(size_t)g * a.N + c] = 0.0;
  // Transient: the cell solves use sigma_t + inv_v / dt (DiamondDifference3D.h:25-26).
  a.sig[c] = tr ? a.sigt[(size_t)m * G + g] + a.inv_v[(size_t)m * G + g] / a.dt
                : a.sigt[(size_t)m * G + g];
}

// stepDNPs (TransportSolver3D.cpp:230-260): with the step's flux, per delayed
// group S = sum_g nuSf_g phi_g beta_gd and C = S / lambda (1 - e^{-lambda dt}) +
// C e^{-lambda dt} (e^{-lambda dt} precomputed on the host with the C library's
// exp); groups past the material's count stay zero.
__global__ void
step_dnps_kernel(IterArgs a)
{
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  const int m = a.mat[c];
  const int nd = a.nd[m];
  for (int d = 0; d < nd; ++d)
  {
    double src = 0.0;
    for (int g = 0; g < a.G; ++g)
      src += a.prod[(size_t)m * a.G + g] * a.phi[(size_t)g * a.N + c] *
             a.beta[((size_t)m * a.G + g) * DOMA_MAX_DELAYED + d];
    const double lam = a.lambda[(size_t)m * DOMA_MAX_DELAYED + d];
    const double e = a.decay[(size_t)m * DOMA_MAX_DELAYED + d];
    const size_t k = (size_t)d * a.N + c;
    a.dnp[k] = (src / lam) * (1.0 - e) + a.dnp[k] * e;
  }
}

// Sum of the direction-set partials of cell c in set order.  Loads are issued
// eight at a time ahead of the additions so that many sets (the 2D row schedule
// writes one per direction) cost a few memory latencies, not one per set; the
// additions keep the set order, so the sum is the same as a plain loop's.
__device__ __forceinline__ double
sum_parts(const double * part, int nsets, size_t N, int c)
{
#pragma clang fp contract(off)
  double d = 0.0;
  int s = 0;
  for (; s + 8 <= nsets; s += 8)
  {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = part[(size_t)(s + u) * N + c];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      d += v[u];
  }
  for (; s < nsets; ++s)
    d += part[(size_t)s * N + c];
  return d;
}

// Sum of the direction-set partials, the scattering residual terms and the
// scattering-source update, fused.  2D/3D increment form: TransportSolver3D.cpp:
// 531-563; 1D standard form: TransportSolver1D.cpp:802-838.  Block `block` of
// `nblocks` walks a fixed strided set of cells.
__device__ __forceinline__ void
finish_body(const IterArgs & a, int block, int nblocks)
{
#pragma clang fp contract(off)
  if (a.si_state && a.si_state[0])
    return;
  double diff = 0.0, total = 0.0;
  const int stride = nblocks * RED_THREADS;
  for (int c = block * RED_THREADS + threadIdx.x; c < a.N; c += stride)
  {
    double d = 0.0;
    if (a.dsa)
      d = a.dsa[c];
    else
      d = sum_parts(a.part, a.ndirsets, (size_t)a.N, c);
    // After a learned correction e the sweep's source was S (d_last + e); the
    // increment from the corrected flux is K (d_last + e) - e.
    if (a.corr)
      d -= a.corr[c];
    if (a.dcur)
      a.dcur[c] = d;
    const double v = measure(a, c);
    const size_t pc = (size_t)a.g * a.N + c;
    const double t = a.phi[pc];
    const int m = a.mat[c];
    const double sgg = a.sigs[((size_t)m * a.G + a.g) * a.G + a.g];
    if (a.dims == 1)
    {
      diff = diff + (d - t) * (d - t) * v;
      total = total + d * d * v;
      a.phi[pc] = d;
      double q = a.qmg[c];
      q += 0.5 * sgg * d;
      a.q[c] = q;
    }
    else
    {
      diff = diff + d * d * v;
      total = total + (t + d) * (t + d) * v;
      a.phi[pc] = t + d;
      a.q[c] = norm_term(a.dims, sgg, d);
    }
  }
  block_sum2(diff, total);
  if (threadIdx.x == 0)
  {
    a.red[2 * block + 0] = diff;
    a.red[2 * block + 1] = total;
  }
}

__global__ __launch_bounds__(RED_THREADS) void
finish_kernel(IterArgs a)
{
  finish_body(a, blockIdx.x, gridDim.x);
}

// Final fixed-order sum of block partials; res[0] = residual as the reference
// forms it (0 when the squared norm is <= 1e-8, TransportSolver3D.cpp:562).
__device__ __forceinline__ void
residual_body(const IterArgs & a, int nblocks)
{
  if (a.si_state && a.si_state[0])
    return;
  double diff = 0.0, total = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += RED_THREADS)
  {
    diff += a.red[2 * b + 0];
    total += a.red[2 * b + 1];
  }
  block_sum2(diff, total);
  if (threadIdx.x == 0)
  {
    const double r = total > 1e-8 ? sqrt(diff) / sqrt(total) : 0.0;
    a.res[0] = r;
    a.res[1] = diff;
    a.res[2] = total;
    a.res[3] = a.err ? static_cast<double>(*a.err) : 0.0;
    if (a.si_state)
    {
      // The loop condition of sourceIteration (TransportSolver3D.cpp:481).
      const int it = ++a.si_state[1];
      if (!(a.tol < r) || it >= a.max_it || a.res[3] != 0.0)
        a.si_state[0] = 1;
      // A copy of the loop state beside the residual, so that the host reads
      // both with one transfer.
      a.res[6] = static_cast<double>(a.si_state[0]);
      a.res[7] = static_cast<double>(it);
    }
  }
}

__global__ __launch_bounds__(RED_THREADS) void
residual_final_kernel(IterArgs a, int nblocks)
{
  residual_body(a, nblocks);
}

// Batched finish: block b works for problem slot map[b].x as its block map[b].y
// of map[b].z; the problems' own block counts keep their reductions identical to
// the unbatched ones.
__global__ __launch_bounds__(RED_THREADS) void
batch_finish_kernel(const IterArgs * args, const int4 * map)
{
  const int4 m = map[blockIdx.x];
  const IterArgs a = args[m.x];
  finish_body(a, m.y, m.z);
}

// Batched residual: block b reduces problem slot b.
__global__ __launch_bounds__(RED_THREADS) void
batch_residual_kernel(const IterArgs * args, const int * nblocks)
{
  const IterArgs a = args[blockIdx.x];
  residual_body(a, nblocks[blockIdx.x]);
}

// Learned correction of group g: keep phi_g, then phi_g += e and the next
// source is S (d_last + e).
__global__ void
apply_correction_kernel(IterArgs a, double * keep)
{
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  const int m = a.mat[c];
  const double sgg = a.sigs[((size_t)m * a.G + a.g) * a.G + a.g];
  const double e = a.corr[c];
  const size_t pc = (size_t)a.g * a.N + c;
  keep[c] = a.phi[pc];
  a.phi[pc] += e;
  a.q[c] = norm_term(a.dims, sgg, a.dlast[c] + e);
}

// Undo a rejected correction: restore the kept flux exactly and the source S d_last.
__global__ void
undo_correction_kernel(IterArgs a, const double * keep)
{
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  const int m = a.mat[c];
  const double sgg = a.sigs[((size_t)m * a.G + a.g) * a.G + a.g];
  a.phi[(size_t)a.g * a.N + c] = keep[c];
  a.q[c] = norm_term(a.dims, sgg, a.dlast[c]);
}

// Warm start of group g from the previous multigroup iterate (prev_g): phi_g =
// prev_g and the first sweep's source gains S prev_g; that sweep's increment is
// then its output minus prev_g (finish with corr = prev_g).
__global__ void
warm_start_kernel(IterArgs a)
{
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  const int m = a.mat[c];
  const double sgg = a.sigs[((size_t)m * a.G + a.g) * a.G + a.g];
  const size_t pc = (size_t)a.g * a.N + c;
  const double f = a.prev[pc];
  a.phi[pc] = f;
  a.q[c] += norm_term(a.dims, sgg, f);
}

// Training-data snapshot after a source iteration: the latest increment (the SI
// residual field, FP32) and the accumulated scalar flux of the group (FP64).
__global__ void
capture_kernel(IterArgs a, float * res_out, double * phi_out)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  const double d = sum_parts(a.part, a.ndirsets, (size_t)a.N, c);
  res_out[c] = static_cast<float>(d);
  if (phi_out)
    phi_out[c] = a.phi[(size_t)a.g * a.N + c];
}

// Targets of the stored snapshots once the group's iteration has converged:
// target_k = phi_converged - phi_k, in FP32.
__global__ void
target_kernel(const double * phi_g, const double * snaps, float * targets, int N, int count)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * count)
    return;
  targets[i] = static_cast<float>(phi_g[i % (size_t)N] - snaps[i]);
}

__global__ void
reduce_parts_kernel(IterArgs a)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  a.dphi[c] = sum_parts(a.part, a.ndirsets, (size_t)a.N, c);
}

// computeMGFluxResidual terms (TransportSolver3D.cpp:318-337).
__global__ __launch_bounds__(RED_THREADS) void
mg_residual_kernel(IterArgs a)
{
#pragma clang fp contract(off)
  double diff = 0.0, total = 0.0;
  const int stride = gridDim.x * RED_THREADS;
  for (int c = blockIdx.x * RED_THREADS + threadIdx.x; c < a.N; c += stride)
  {
    const double v = measure(a, c);
    for (int g = 0; g < a.G; ++g)
    {
      const double f = a.phi[(size_t)g * a.N + c];
      const double d = f - a.prev[(size_t)g * a.N + c];
      diff = diff + d * d * v;
      total = total + f * f * v;
    }
  }
  block_sum2(diff, total);
  if (threadIdx.x == 0)
  {
    a.red[2 * blockIdx.x + 0] = diff;
    a.red[2 * blockIdx.x + 1] = total;
  }
}

// updateEigenvalue sums (TransportSolver3D.cpp:432-456; 2D :432-456; 1D
// TransportSolver1D.cpp:468-493) over materials with nu_sigma_f: per cell and
// group, num += V nuSf phi and den += V nuSf prev (/ k in 2D/3D).
__global__ __launch_bounds__(RED_THREADS) void
fission_rate_kernel(IterArgs a)
{
#pragma clang fp contract(off)
  double num = 0.0, den = 0.0;
  const int stride = gridDim.x * RED_THREADS;
  for (int c = blockIdx.x * RED_THREADS + threadIdx.x; c < a.N; c += stride)
  {
    const int m = a.mat[c];
    if (!a.has_prod[m])
      continue;
    const double v = measure(a, c);
    for (int g = 0; g < a.G; ++g)
    {
      const double vp = v * a.prod[(size_t)m * a.G + g];
      num = num + vp * a.phi[(size_t)g * a.N + c];
      if (a.dims == 1)
        den = den + vp * a.prev[(size_t)g * a.N + c];
      else
        den = den + vp * a.prev[(size_t)g * a.N + c] / a.k;
    }
  }
  block_sum2(num, den);
  if (threadIdx.x == 0)
  {
    a.red[2 * blockIdx.x + 0] = num;
    a.red[2 * blockIdx.x + 1] = den;
  }
}

__global__ __launch_bounds__(RED_THREADS) void
rate_final_kernel(IterArgs a, int nblocks)
{
  double num = 0.0, den = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += RED_THREADS)
  {
    num += a.red[2 * b + 0];
    den += a.red[2 * b + 1];
  }
  block_sum2(num, den);
  if (threadIdx.x == 0)
  {
    a.res[4] = num;
    a.res[5] = den;
  }
}

// syntheticAcceleration (TransportSolver1D.cpp:587-800; Alcouffe, NSE 64 (1977)
// 344, with Larsen's boundary conditions, NSE 82 (1982) 47): one work-group.
// Moments of the sweep from the direction partials part[t][c] = w_t psi_t(c)
// (t < dirs: mu > 0, swept left to right): phi = sum_t part, J = sum_t mu_t part,
// edge scalar fluxes from the cell-centre fluxes tallied onto downwind edges
// (DiamondDifference1D.h:41-43, CartesianCell1D.cpp:116-120).  Rows m = 1 .. N-1
// couple cells m-1 and m: fixed-source mode uses the source correction (Eq. 22,
// 23), eigenvalue mode the modified diffusion coefficients D = -h J / (sfR - sfL)
// unless one is <= 0, then the removal correction.  The (N+1)-unknown system is
// solved by Gaussian elimination with partial pivoting (the tridiagonal form of
// LAPACK's dgtsv) in one lane; the cell flux is the average of its two edges.
__global__ __launch_bounds__(RED_THREADS) void
dsa1d_kernel(IterArgs a, DsaArgs d)
{
#pragma clang fp contract(off)
  if (a.si_state && a.si_state[0])
    return;
  const int N = a.N, n = N + 1, G = a.G, g = a.g, D = d.dirs;
  double * phi = d.work;
  double * cur = phi + N;
  double * sfr = cur + N;
  double * sfl = sfr + N;
  double * lo = sfl + N;   // A[m][m-1]
  double * di = lo + n;    // A[m][m]
  double * up = di + n;    // A[m][m+1]
  double * up2 = up + n;   // second super-diagonal created by row interchanges
  double * rhs = up2 + n;
  double * x = rhs + n;
  for (int c = threadIdx.x; c < N; c += RED_THREADS)
  {
    // Edge e collects w psi of cell e-1 for mu > 0 and of cell e for mu < 0
    // (CartesianCell1D.cpp:116-120); the cell's right edge is c+1, its left c.
    double f = 0.0, j = 0.0, r = 0.0, l = 0.0;
    for (int t = 0; t < 2 * D; ++t)
    {
      const double v = a.part[(size_t)t * N + c];
      const double mu = t < D ? d.dirtab[4 * t] : -d.dirtab[4 * t];
      f += v;
      j += mu * v;
      if (t < D)
      {
        r += v;
        if (c > 0)
          l += a.part[(size_t)t * N + c - 1];
      }
      else
      {
        l += v;
        if (c + 1 < N)
          r += a.part[(size_t)t * N + c + 1];
      }
    }
    phi[c] = f;
    cur[c] = j;
    sfr[c] = r;
    sfl[c] = l;
  }
  __syncthreads();
  auto tot = [&](int c) { return a.sigt[(size_t)a.mat[c] * G + g]; };
  auto self = [&](int c) { return a.sigs[((size_t)a.mat[c] * G + g) * G + g]; };
  // sigma_r at the edge between cells i and i + 1 (Eq. 23a).
  auto sig_r = [&](int i) {
    double sr = (tot(i) - self(i)) * a.lx[i] * phi[i];
    sr += (tot(i + 1) - self(i + 1)) * a.lx[i + 1] * phi[i + 1];
    sr *= 0.5 / sfr[i];
    return sr;
  };
  bool removal = false;
  if (a.eigen)
  {
    int bad = 0;
    for (int i = threadIdx.x; i < N - 1; i += RED_THREADS)
    {
      double Di = -1.0 * a.lx[i] * cur[i];
      Di /= (sfr[i] - sfl[i]);
      double Di1 = -1.0 * a.lx[i + 1] * cur[i + 1];
      Di1 /= (sfr[i + 1] - sfl[i + 1]);
      if (Di <= 0.0 || Di1 <= 0.0)
        bad = 1;
      else
      {
        const int m = i + 1;
        lo[m] = -1.0 * Di / a.lx[i];
        di[m] = (Di / a.lx[i]) + (Di1 / a.lx[i + 1]) + sig_r(i);
        up[m] = -1.0 * Di1 / a.lx[i + 1];
        rhs[m] = a.qmg[i] * a.lx[i] + a.qmg[i + 1] * a.lx[i + 1];
      }
    }
    removal = __syncthreads_or(bad) != 0;
  }
  if (!a.eigen || removal)
    for (int i = threadIdx.x; i < N - 1; i += RED_THREADS)
    {
      const int m = i + 1;
      const double Dh = 1.0 / 3.0 / tot(i) / a.lx[i];
      const double Dh1 = 1.0 / 3.0 / tot(i + 1) / a.lx[i + 1];
      double corr = cur[i + 1] - cur[i];
      corr += Dh1 * (sfr[i + 1] - sfl[i + 1]);
      corr -= Dh * (sfr[i] - sfl[i]);
      const double qqh = a.qmg[i] * a.lx[i] + a.qmg[i + 1] * a.lx[i + 1];
      lo[m] = -1.0 * Dh;
      up[m] = -1.0 * Dh1;
      if (a.eigen)  // removal correction
      {
        corr /= sfr[i];
        di[m] = Dh + Dh1 + sig_r(i) + corr;
        rhs[m] = qqh;
      }
      else  // source correction
      {
        di[m] = Dh + Dh1 + sig_r(i);
        rhs[m] = qqh - corr;
      }
    }
  if (threadIdx.x == 0)
  {
    // Boundary rows (Larsen Eq. 34, 37).
    const double hd0 = 0.5 / 3.0 / tot(0) / a.lx[0];
    const double es0 = 0.125 * (tot(0) - self(0)) * a.lx[0];
    lo[0] = 0.0;
    di[0] = d.beta_left + hd0 + es0;
    up[0] = -1.0 * hd0 + es0;
    rhs[0] = -1.0 * a.qmg[0];
    const int I = N - 1;
    const double hdI = 0.5 / 3.0 / tot(I) / a.lx[I];
    const double esI = 0.125 * (tot(I) - self(I)) * a.lx[I];
    di[N] = d.beta_right - hdI - esI;
    lo[N] = hdI - esI;
    up[N] = 0.0;
    rhs[N] = a.qmg[I];
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    // Forward elimination with partial pivoting between rows i and i + 1.
    for (int i = 0; i < n - 1; ++i)
    {
      up2[i] = 0.0;
      if (fabs(di[i]) >= fabs(lo[i + 1]))
      {
        const double f = lo[i + 1] / di[i];
        di[i + 1] = di[i + 1] - f * up[i];
        rhs[i + 1] = rhs[i + 1] - f * rhs[i];
      }
      else
      {
        // Rows i and i + 1 swap: row i becomes [lo[i+1], di[i+1], up[i+1]].
        const double f = di[i] / lo[i + 1];
        di[i] = lo[i + 1];
        const double t = di[i + 1];
        di[i + 1] = up[i] - f * t;
        if (i < n - 2)
        {
          up2[i] = up[i + 1];
          up[i + 1] = -f * up2[i];
        }
        up[i] = t;
        const double r = rhs[i];
        rhs[i] = rhs[i + 1];
        rhs[i + 1] = r - f * rhs[i + 1];
      }
    }
    x[n - 1] = rhs[n - 1] / di[n - 1];
    if (n > 1)
      x[n - 2] = (rhs[n - 2] - up[n - 2] * x[n - 1]) / di[n - 2];
    for (int i = n - 3; i >= 0; --i)
      x[i] = (rhs[i] - up[i] * x[i + 1] - up2[i] * x[i + 2]) / di[i];
  }
  __syncthreads();
  for (int c = threadIdx.x; c < N; c += RED_THREADS)
    d.out[c] = 0.5 * (x[c] + x[c + 1]);
}

__global__ void
fill_i32_kernel(int * p, size_t n, int v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

__global__ void
fill_kernel(double * p, size_t n, double v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// Start of a single-group source iteration from a given multigroup source:
// phi_g = 0 and q (2D/3D) or qmg (1D) = qmg_dev, with the group's cross sections.
__global__ void
set_q_kernel(IterArgs a, const double * qmg_dev)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  if (a.dims == 1)
    a.qmg[c] = qmg_dev[c];
  else
    a.q[c] = qmg_dev[c];
  a.phi[(size_t)a.g * a.N + c] = 0.0;
  a.sig[c] = a.sigt[(size_t)a.mat[c] * a.G + a.g];
}

// 1D: sourceIteration applies updateScatteringSource once before the first
// sweep (TransportSolver1D.cpp:500): phi_g = dphi (= 0), q = qmg.
__global__ void
q_from_qmg_kernel(IterArgs a)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  a.q[c] = a.qmg[c];
  a.phi[(size_t)a.g * a.N + c] = 0.0;
}

// Gather of the swept group's total cross section for the plain sweep entry point.
__global__ void
gather_sig_kernel(IterArgs a)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N)
    return;
  a.sig[c] = a.sigt[(size_t)a.mat[c] * a.G + a.g];
}

inline unsigned
blocks_for(size_t n, int t)
{
  return (unsigned)((n + t - 1) / t);
}
} // namespace

// Blocks of the fixed-order reductions over cells: one cell per thread up to
// 2048 blocks (at 64^3, 1024 blocks keep 4 waves per SIMD of the finish
// kernel's loads in flight; 256 blocks of 4 cells per thread left it latency
// bound).
int
reduction_blocks(int N)
{
  int b = (N + RED_THREADS - 1) / RED_THREADS;
  if (b < 1)
    b = 1;
  if (b > 2048)
    b = 2048;
  return b;
}

int
launch_mg_source(const IterArgs & a, hipStream_t s)
{
  hipLaunchKernelGGL(mg_source_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
  if (a.dims == 1)
    hipLaunchKernelGGL(q_from_qmg_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
  return DOMA_OK;
}

int
launch_set_q(const IterArgs & a, const double * qmg_dev, hipStream_t s)
{
  hipLaunchKernelGGL(set_q_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a, qmg_dev);
  if (a.dims == 1)
    hipLaunchKernelGGL(q_from_qmg_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
  return DOMA_OK;
}

int
launch_gather_sig(const IterArgs & a, hipStream_t s)
{
  hipLaunchKernelGGL(gather_sig_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
  return DOMA_OK;
}

int
launch_finish(const IterArgs & a, int nblocks, hipStream_t s)
{
  hipLaunchKernelGGL(finish_kernel, dim3(nblocks), dim3(RED_THREADS), 0, s, a);
  return DOMA_OK;
}

int
launch_residual_final(const IterArgs & a, int nblocks, hipStream_t s)
{
  hipLaunchKernelGGL(residual_final_kernel, dim3(1), dim3(RED_THREADS), 0, s, a, nblocks);
  return DOMA_OK;
}

int
launch_reduce_parts(const IterArgs & a, hipStream_t s)
{
  hipLaunchKernelGGL(reduce_parts_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
  return DOMA_OK;
}

int
launch_mg_residual(const IterArgs & a, int nblocks, hipStream_t s)
{
  hipLaunchKernelGGL(mg_residual_kernel, dim3(nblocks), dim3(RED_THREADS), 0, s, a);
  hipLaunchKernelGGL(residual_final_kernel, dim3(1), dim3(RED_THREADS), 0, s, a, nblocks);
  return DOMA_OK;
}

int
launch_fission_rates(const IterArgs & a, int nblocks, hipStream_t s)
{
  hipLaunchKernelGGL(fission_rate_kernel, dim3(nblocks), dim3(RED_THREADS), 0, s, a);
  hipLaunchKernelGGL(rate_final_kernel, dim3(1), dim3(RED_THREADS), 0, s, a, nblocks);
  return DOMA_OK;
}

int
launch_step_dnps(const IterArgs & a, hipStream_t s)
{
  hipLaunchKernelGGL(step_dnps_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
  return DOMA_OK;
}

int
launch_dsa1d(const IterArgs & a, const DsaArgs & d, hipStream_t s)
{
  hipLaunchKernelGGL(dsa1d_kernel, dim3(1), dim3(RED_THREADS), 0, s, a, d);
  return DOMA_OK;
}

int
launch_batch_finish(const IterArgs * d_args, const int4 * d_map, int nblocks, hipStream_t s)
{
  if (nblocks > 0)
    hipLaunchKernelGGL(batch_finish_kernel, dim3(nblocks), dim3(RED_THREADS), 0, s, d_args, d_map);
  return DOMA_OK;
}

int
launch_batch_residual(const IterArgs * d_args, const int * d_nblocks, int nprob, hipStream_t s)
{
  if (nprob > 0)
    hipLaunchKernelGGL(batch_residual_kernel, dim3(nprob), dim3(RED_THREADS), 0, s, d_args,
                       d_nblocks);
  return DOMA_OK;
}

int
launch_capture(const IterArgs & a, float * res_out, double * phi_out, hipStream_t s)
{
  hipLaunchKernelGGL(capture_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a, res_out,
                     phi_out);
  return DOMA_OK;
}

int
launch_targets(const double * phi_g, const double * snaps, float * targets, int N, int count,
               hipStream_t s)
{
  if (count > 0)
    hipLaunchKernelGGL(target_kernel, dim3(blocks_for((size_t)N * count, 256)), dim3(256), 0, s,
                       phi_g, snaps, targets, N, count);
  return DOMA_OK;
}

int
launch_apply_correction(const IterArgs & a, double * keep, hipStream_t s)
{
  hipLaunchKernelGGL(apply_correction_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a,
                     keep);
  return DOMA_OK;
}

int
launch_undo_correction(const IterArgs & a, const double * keep, hipStream_t s)
{
  hipLaunchKernelGGL(undo_correction_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a,
                     keep);
  return DOMA_OK;
}

int
launch_warm_start(const IterArgs & a, hipStream_t s)
{
  hipLaunchKernelGGL(warm_start_kernel, dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
  return DOMA_OK;
}

int
launch_fill_i32(int * p, size_t n, int v, hipStream_t s)
{
  if (n == 0)
    return DOMA_OK;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
  return DOMA_OK;
}

int
launch_fill(double * p, size_t n, double v, hipStream_t s)
{
  if (n == 0)
    return DOMA_OK;
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
  return DOMA_OK;
}
} // namespace doma
