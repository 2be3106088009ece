// sweep_rows.hip -- 2D sweep with one lane per mesh row (schedule 2).
//
// A work-group sweeps one (quadrant, direction) over the whole mesh.  Lane jj
// owns row jj (counted in the sweep's y order) and at step s solves the cell in
// column ii = s - jj, so every step solves one antidiagonal and a sweep is
// nx + ny - 1 dependent cell solves (TransportSolver2D.cpp:635-725 visits the
// same cells row by row).  The incoming x face stays in a register; the
// incoming y face is the one the lane below produced a step earlier, passed by a
// wave shift and, across waves, through LDS.  The inputs of a lane's next eight
// cells are loaded eight steps ahead, and the column widths sit in LDS, so a
// step costs a cell solve, a shift and a barrier rather than a global-memory
// round trip.  Each work-group writes its direction's w psi into its own
// partial-flux plane, summed in direction order by the finish kernel.
#include <hip/hip_runtime.h>

#include "cellsolve.h"
#include "internal.h"

namespace doma
{
namespace
{
constexpr int kAhead = 8;  // steps between a cell's loads and its solve

template <typename R, int SCHEME, bool EXACT>
__device__ __forceinline__ R
cell2d(double q, double sg, double lx, double ly, R amu, R aeta, R axi, R & x, R & y)
{
  if constexpr (EXACT)
  {
    if constexpr (SCHEME == DOMA_SCHEME_TW_DD)
      return twdd2d_exact<R>(R(q), R(sg), R(lx), R(ly), amu, aeta, axi, x, y);
    else
      return dd2d_exact<R>(R(q), R(sg), R(lx), R(ly), amu, aeta, x, y);
  }
  else if constexpr (SCHEME == DOMA_SCHEME_TW_DD)
  {
    // The brick kernel's fast TW-DD inputs: q V, sigma V, A = ly, B = lx.
    const double V = lx * ly;
    return twdd2d_fast<R>(R(q * V), R(sg * V), amu * R(ly), aeta * R(lx), axi, x, y);
  }
  else
  {
    R z = R(0);
    return dd3d_fast<R>(R(q), R(sg), R(2) * amu * R(1.0 / lx), R(2) * aeta * R(1.0 / ly), R(0), x,
                        y, z);
  }
}

template <typename R, int SCHEME, bool EXACT>
__global__ __launch_bounds__(kRowsMaxThreads) void
sweep2d_rows_kernel(Rows2DArgs a)
{
  if (a.skip && *a.skip)
    return;
  extern __shared__ __align__(16) unsigned char smem[];
  double * s_lx = reinterpret_cast<double *>(smem);          // [nx] in sweep order
  R * s_y = reinterpret_cast<R *>(smem + 8 * (size_t)a.nx);  // [2][waves] wave-edge y faces
  const int slot = blockIdx.x / a.dirs;  // quadrant: bit 2 = -mu, bit 1 = -eta
  const bool rev_x = (slot & 2) != 0, rev_y = (slot & 1) != 0;
  const double * d = a.dirtab + 4 * blockIdx.x;
  const R amu = R(d[0]), aeta = R(d[1]), axi = R(d[2]);
  const double w = d[3];
  for (int ii = threadIdx.x; ii < a.nx; ii += blockDim.x)
    s_lx[ii] = a.lx[rev_x ? a.nx - 1 - ii : ii];
  const int jj = threadIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool row = jj < a.ny;
  const int j = rev_y ? a.ny - 1 - jj : jj;
  const double ly = row ? a.ly[j] : 1.0;
  const double * qrow = a.q + (size_t)(row ? j : 0) * a.nx;
  const double * srow = a.sig + (size_t)(row ? j : 0) * a.nx;
  double * prow = a.part + (size_t)blockIdx.x * a.N + (size_t)(row ? j : 0) * a.nx;
  auto col = [&](int ii) { return rev_x ? a.nx - 1 - ii : ii; };
  auto clampc = [&](int ii) { return col(ii < 0 ? 0 : (ii >= a.nx ? a.nx - 1 : ii)); };
  // Inputs of the cells of steps s0 .. s0+7 (cur) and of the next eight (nxt).
  double cq[kAhead], cs[kAhead], nq[kAhead], ns[kAhead];
#pragma unroll
  for (int u = 0; u < kAhead; ++u)
  {
    const int c = clampc(u - jj);
    cq[u] = qrow[c];
    cs[u] = srow[c];
  }
  __syncthreads();
  R x = R(0), yo = R(0);
  const int nsteps = a.nx + a.ny - 1;
  for (int s0 = 0; s0 < nsteps; s0 += kAhead)
  {
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
    {
      const int c = clampc(s0 + kAhead + u - jj);
      nq[u] = qrow[c];
      ns[u] = srow[c];
    }
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
    {
      const int s = s0 + u;
      if (s >= nsteps)
        break;
      // y face from the lane below, produced at step s - 1.
      R yin = __shfl_up(yo, 1, 64);
      if (lane == 0)
        yin = wave > 0 && s > 0 ? s_y[((s - 1) & 1) * 16 + wave - 1] : R(0);
      if (jj == 0)
        yin = R(0);
      const int ii = s - jj;
      if (row && ii >= 0 && ii < a.nx)
      {
        R y = yin;
        const R psi = cell2d<R, SCHEME, EXACT>(cq[u], cs[u], s_lx[ii], ly, amu, aeta, axi, x, y);
        yo = y;
        prow[col(ii)] = w * static_cast<double>(psi);
      }
      if (lane == 63)
        s_y[(s & 1) * 16 + wave] = yo;
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
    {
      cq[u] = nq[u];
      cs[u] = ns[u];
    }
  }
}

template <typename R, int SCHEME, bool EXACT>
int
launch_rows(const Rows2DArgs & a, hipStream_t s)
{
  const int threads = ((a.ny + 63) / 64) * 64;
  const size_t lds = 8 * (size_t)a.nx + 2 * 16 * sizeof(R);
  hipLaunchKernelGGL((sweep2d_rows_kernel<R, SCHEME, EXACT>), dim3(4 * a.dirs), dim3(threads),
                     lds, s, a);
  return DOMA_OK;
}
} // namespace

bool
rows2d_fits(int nx, int ny)
{
  return ny >= 1 && ny <= kRowsMaxThreads && nx >= 1 && 8 * (size_t)nx + 256 <= 64 * 1024;
}

int
launch_sweep2d_rows(const Rows2DArgs & a, int scheme, int precision, int exact, hipStream_t s)
{
  if (!rows2d_fits(a.nx, a.ny))
    return fail(DOMA_ERR_INVALID, "row-pipelined 2D sweep needs ny <= " +
                                      std::to_string(kRowsMaxThreads) + " and nx <= 8160");
  const bool f32 = precision == DOMA_FP32_SWEEP;
  if (scheme == DOMA_SCHEME_TW_DD)
  {
    if (exact)
      return f32 ? launch_rows<float, DOMA_SCHEME_TW_DD, true>(a, s)
                 : launch_rows<double, DOMA_SCHEME_TW_DD, true>(a, s);
    return f32 ? launch_rows<float, DOMA_SCHEME_TW_DD, false>(a, s)
               : launch_rows<double, DOMA_SCHEME_TW_DD, false>(a, s);
  }
  if (exact)
    return f32 ? launch_rows<float, DOMA_SCHEME_DD, true>(a, s)
               : launch_rows<double, DOMA_SCHEME_DD, true>(a, s);
  return f32 ? launch_rows<float, DOMA_SCHEME_DD, false>(a, s)
             : launch_rows<double, DOMA_SCHEME_DD, false>(a, s);
}
} // namespace doma
