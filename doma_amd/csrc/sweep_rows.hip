// sweep_rows.hip -- 2D sweep with one lane per mesh row (schedule 2).
//
// A work-group sweeps one (quadrant, direction) over the whole mesh.  Lane jj
// owns row jj (counted in the sweep's y order) and at step s solves the cell in
// column ii = s - jj, so every step solves one antidiagonal and a sweep is
// nx + ny - 1 dependent cell solves (TransportSolver2D.cpp:635-725 visits the
// same cells row by row).  The incoming x face stays in a register; the
// incoming y face is the one the lane below produced a step earlier, passed by a
// wave shift, and across waves through an LDS column of faces with a ready
// count, so waves run their own step windows without a block barrier (a wave
// only waits for the one below it, which never waits for it).  The inputs of a
// lane's next eight cells are loaded eight steps ahead and the column widths sit
// in LDS, so a step costs a cell solve and a shift rather than a global-memory
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

// Lane i receives lane i - 1's value (lane 0 keeps its own): the DPP wave shift
// right by one (wave_shr:1, GFX9 DPP control 0x138), a VALU move instead of an
// LDS permute on the step-to-step dependency chain.
__device__ __forceinline__ int
wave_shr1(int v)
{
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false);
}

__device__ __forceinline__ double
wave_shr1(double v)
{
  const int lo = wave_shr1(__double2loint(v));
  const int hi = wave_shr1(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float
wave_shr1(float v)
{
  return __int_as_float(wave_shr1(__float_as_int(v)));
}

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
  const int nwaves = (int)(blockDim.x >> 6);
  double * s_lx = reinterpret_cast<double *>(smem);  // [nx] in sweep order
  // y faces a wave's top lane hands to the next wave, by column, and how many
  // columns of them are ready.
  R * s_y = reinterpret_cast<R *>(smem + 8 * (size_t)a.nx);
  int * s_ready = reinterpret_cast<int *>(smem + 8 * (size_t)a.nx +
                                          sizeof(R) * (size_t)a.nx * (nwaves - 1));
  const int slot = blockIdx.x / a.dirs;  // quadrant: bit 2 = -mu, bit 1 = -eta
  const bool rev_x = (slot & 2) != 0, rev_y = (slot & 1) != 0;
  const double * d = a.dirtab + 4 * blockIdx.x;
  const R amu = R(d[0]), aeta = R(d[1]), axi = R(d[2]);
  const double w = d[3];
  for (int ii = threadIdx.x; ii < a.nx; ii += blockDim.x)
    s_lx[ii] = a.lx[rev_x ? a.nx - 1 - ii : ii];
  if (threadIdx.x < 16)
    s_ready[threadIdx.x] = 0;
  const int jj = threadIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool row = jj < a.ny;
  const int j = rev_y ? a.ny - 1 - jj : jj;
  const double ly = row ? a.ly[j] : 1.0;
  const double * qrow = a.q + (size_t)(row ? j : 0) * a.nx;
  const double * srow = a.sig + (size_t)(row ? j : 0) * a.nx;
  double * prow = a.part + (size_t)blockIdx.x * a.N + (size_t)(row ? j : 0) * a.nx;
  auto col = [&](int ii) { return rev_x ? a.nx - 1 - ii : ii; };
  auto clampc = [&](int ii) { return col(ii < 0 ? 0 : (ii >= a.nx ? a.nx - 1 : ii)); };
  // A wave works the steps in which one of its rows has a column to solve:
  // s = 64 wave .. 64 wave + (its rows - 1) + nx - 1, ahead of the waves above.
  const int wave_rows = min(64, a.ny - 64 * wave);
  const int s_first = 64 * wave, s_last = 64 * wave + wave_rows - 1 + a.nx - 1;
  double cq[kAhead], cs[kAhead], nq[kAhead], ns[kAhead];
#pragma unroll
  for (int u = 0; u < kAhead; ++u)
  {
    const int c = clampc(s_first + u - jj);
    cq[u] = qrow[c];
    cs[u] = srow[c];
  }
  __syncthreads();
  const R * y_below = s_y + (size_t)(wave - 1) * a.nx;  // wave > 0 only
  R * y_above = s_y + (size_t)wave * a.nx;             // wave < nwaves - 1 only
  // seen: columns of the wave below known to be ready (uniform).  The wave below
  // publishes its count once per group of steps, so the count is read only when
  // this wave runs past the last value it saw.
  int published = 0, seen = 0;
  R x = R(0), yo = R(0);
  // Eight steps on the inputs in (uq, us) while the next eight cells' inputs
  // load into (lq, ls); two register sets alternate, so no copy waits on them.
  auto group = [&](int s0, const double (&uq)[kAhead], const double (&us)[kAhead],
                   double (&lq)[kAhead], double (&ls)[kAhead]) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
    {
      const int c = clampc(s0 + kAhead + u - jj);
      lq[u] = qrow[c];
      ls[u] = srow[c];
    }
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
    {
      const int s = s0 + u;
      if (s > s_last)
        break;
      // y face from the lane below, produced at step s - 1; lane 0 takes the
      // wave below's top-row face of the same column, once it is ready (the
      // wave below never waits on this one, so the wait ends).
      R yin = wave_shr1(yo);
      const int ii0 = s - 64 * wave;  // lane 0's column (uniform)
      if (wave > 0 && ii0 >= 0 && ii0 < a.nx)
      {
        if (ii0 >= seen)
        {
          seen = __hip_atomic_load(s_ready + wave - 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          if (seen <= ii0)
          {
            // Bounded by a wall-clock deadline, as the dataflow waits are; the
            // sleep leaves issue slots to the producing wave on the same CU.
            const long long t0 = wall_clock64();
            int polls = 0;
            while ((seen = __hip_atomic_load(s_ready + wave - 1, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP)) <= ii0)
            {
              __builtin_amdgcn_s_sleep(1);
              if ((++polls & 15) == 0 && wall_clock64() - t0 > a.wait_ticks)
              {
                if (lane == 0)
                  __hip_atomic_store(a.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                seen = a.nx;  // give up: the error flag reports the stall
                break;
              }
            }
          }
          // Orders every later read of the column after the count that covers it.
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");  // LDS only
        }
        const R yb = y_below[ii0];
        if (lane == 0)
          yin = yb;
      }
      if (jj == 0)
        yin = R(0);
      const int ii = s - jj;
      if (row && ii >= 0 && ii < a.nx)
      {
        R y = yin;
        const R psi = cell2d<R, SCHEME, EXACT>(uq[u], us[u], s_lx[ii], ly, amu, aeta, axi, x, y);
        yo = y;
        prow[col(ii)] = w * static_cast<double>(psi);
      }
      // The top lane leaves its face for the wave above.
      const int ii63 = s - 64 * wave - 63;
      if (wave < nwaves - 1 && ii63 >= 0 && ii63 < a.nx)
      {
        if (lane == 63)
          y_above[ii63] = yo;
        published = ii63 + 1;
      }
    }
    // Once per group: order the group's faces before their count (LDS only: the
    // flux stores to global memory are not waited for), then publish the count.
    // The wave above starts 64 steps behind, so the batching does not delay it.
    if (wave < nwaves - 1 && published > 0)
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      if (lane == 63)
        __hip_atomic_store(s_ready + wave, published, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  };
  for (int s0 = s_first; s0 <= s_last; s0 += 2 * kAhead)
  {
    group(s0, cq, cs, nq, ns);
    if (s0 + kAhead <= s_last)
      group(s0 + kAhead, nq, ns, cq, cs);
  }
}

template <typename R, int SCHEME, bool EXACT>
int
launch_rows(const Rows2DArgs & a, hipStream_t s)
{
  const int threads = ((a.ny + 63) / 64) * 64;
  const size_t lds = rows2d_lds(a.nx, a.ny, sizeof(R));
  hipLaunchKernelGGL((sweep2d_rows_kernel<R, SCHEME, EXACT>), dim3(4 * a.dirs), dim3(threads),
                     lds, s, a);
  return DOMA_OK;
}
} // namespace

// LDS: column widths, the wave-to-wave face columns, 16 ready counters.
size_t
rows2d_lds(int nx, int ny, size_t rs)
{
  const int nwaves = (ny + 63) / 64;
  return 8 * (size_t)nx + rs * (size_t)nx * (nwaves - 1) + 16 * sizeof(int);
}

bool
rows2d_fits(int nx, int ny)
{
  return ny >= 1 && ny <= kRowsMaxThreads && nx >= 1 && rows2d_lds(nx, ny, 8) <= 64 * 1024;
}

int
launch_sweep2d_rows(const Rows2DArgs & a, int scheme, int precision, int exact, hipStream_t s)
{
  if (!rows2d_fits(a.nx, a.ny))
    return fail(DOMA_ERR_INVALID, "row-pipelined 2D sweep needs ny <= " +
                                      std::to_string(kRowsMaxThreads) +
                                      " and nx * ny within its LDS budget");
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
