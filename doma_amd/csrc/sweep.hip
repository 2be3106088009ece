// sweep.hip -- brick-wavefront transport sweep for CDNA4 (gfx950).
//
// Decomposition (2D and 3D):
//   * The grid is cut into B^3 bricks (B^2 tiles in 2D).  For every octant the
//     bricks are visited in the octant's own sweep orientation; a brick depends
//     on its upwind neighbours along x, y and z only.  Vacuum boundaries make the
//     octants independent (TransportSolver3D.cpp:565-607, CartesianCell3D.cpp:
//     31-38), so every octant and direction set is swept at once.
//   * One work-group owns one (brick, octant, direction set) at a time.  Inside
//     the brick the cells are processed by hyperplanes li+lj+lk = p; the cells of
//     a plane are independent and run concurrently.  Threads are (cell slot,
//     direction): DL consecutive lanes carry the directions of one cell, so the
//     per-cell source and cross section are uniform per lane group.
//   * Everything a brick needs is staged in LDS before its first plane: the
//     upwind face fluxes (B^2 x-, y- and z-faces per direction, updated in place),
//     the cell sources and cross sections, the cell widths and the plane table.
//     The plane loop then touches no global memory except one store of the
//     direction-reduced scalar flux per cell; only the brick's inflow/outflow
//     faces cross HBM, through per-direction-set face planes.
//   * The scalar flux of a cell is reduced over the DL direction lanes with a
//     fixed DPP butterfly (plus cross-row shuffles above 16 lanes) and stored once
//     per (cell, direction set); the direction sets are summed in a fixed order
//     afterwards, so results are reproducible run to run (no float atomics).
//
// Two schedules share the brick body:
//   * diagonal: one launch per brick diagonal bi+bj+bk = s;
//   * dataflow: one persistent launch per sweep.  Work-groups take (brick,
//     direction set) tasks from a monotone ticket counter in diagonal order and
//     wait on the epoch flags of the (at most three) upwind bricks.  A task only
//     ever waits on tasks with smaller tickets, which were taken by work-groups
//     that are already running, so the schedule cannot deadlock; every wait is
//     bounded and reports through an error flag instead of spinning forever.
#include <hip/hip_runtime.h>

#include "cellsolve.h"
#include "internal.h"

namespace doma
{
namespace
{
// ------------------------------------------------------------------------- //
// Direction reduction                                                        //
// ------------------------------------------------------------------------- //

template <int CTRL>
__device__ __forceinline__ double
dpp_mov(double v)
{
  const int lo = __double2loint(v);
  const int hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
  const int rhi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(rhi, rlo);
}

// DPP controls (GFX9 encoding): quad_perm(1,0,3,2), quad_perm(2,3,0,1),
// row_half_mirror, row_mirror.
constexpr int DPP_QP_XOR1 = 0xB1;
constexpr int DPP_QP_XOR2 = 0x4E;
constexpr int DPP_HALF_MIRROR = 0x141;
constexpr int DPP_MIRROR = 0x140;

// Fixed-order sum over groups of DL (8..64) consecutive lanes; every lane of a
// group ends with the same bitwise result (each step adds a commutative pair).
template <int DL>
__device__ __forceinline__ double
group_sum(double v)
{
  static_assert(DL >= 8 && DL <= 64, "direction lanes");
  v += dpp_mov<DPP_QP_XOR1>(v);
  v += dpp_mov<DPP_QP_XOR2>(v);
  v += dpp_mov<DPP_HALF_MIRROR>(v);
  if (DL >= 16)
    v += dpp_mov<DPP_MIRROR>(v);
  if (DL >= 32)
    v += __shfl_xor(v, 16, 32);
  if (DL >= 64)
    v += __shfl_xor(v, 32, 64);
  return v;
}

// ------------------------------------------------------------------------- //
// Brick body                                                                 //
// ------------------------------------------------------------------------- //

template <typename R, int DIM, int B, int DL>
struct Layout
{
  static constexpr int FS = DIM == 3 ? B * B : B;          // face slots per axis
  static constexpr int NC = DIM == 3 ? B * B * B : B * B;  // cells per brick
  static constexpr int NPLANES = DIM == 3 ? 3 * B - 2 : 2 * B - 1;
  static constexpr size_t FACE_BYTES = sizeof(R) * FS * DL * (DIM == 3 ? 3 : 2);
  static constexpr size_t CELL_OFF = (FACE_BYTES + 15) / 16 * 16;
  static constexpr size_t GEO_OFF = CELL_OFF + 16 * NC;      // double2 {q, sigma_t}
  static constexpr size_t PT_OFF = GEO_OFF + 8 * 3 * B;      // lx, ly, lz per brick
  static constexpr size_t PO_OFF = PT_OFF + 4 * NC;          // plane table
  static constexpr size_t TASK_OFF = PO_OFF + 4 * (NPLANES + 1);
  static constexpr size_t BYTES = TASK_OFF + 16;
};

template <typename R, int DIM, int B, int DL, int NT>
__device__ __forceinline__ void
load_plane_table(const BrickArgs & a, unsigned char * smem)
{
  using L = Layout<R, DIM, B, DL>;
  uint32_t * pt = reinterpret_cast<uint32_t *>(smem + L::PT_OFF);
  int32_t * po = reinterpret_cast<int32_t *>(smem + L::PO_OFF);
  for (int e = threadIdx.x; e < L::NC; e += NT)
    pt[e] = a.plane_cells[e];
  for (int e = threadIdx.x; e <= L::NPLANES; e += NT)
    po[e] = a.plane_off[e];
}

template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
__device__ __forceinline__ void
sweep_brick(const BrickArgs & a, int bi, int bj, int bk, int ds, unsigned char * smem)
{
  using L = Layout<R, DIM, B, DL>;
  constexpr int CPP = NT / DL;  // cells per pass
  constexpr int FS = L::FS;
  R * sx = reinterpret_cast<R *>(smem);
  R * sy = sx + FS * DL;
  R * sz = sy + FS * DL;  // 3D only
  double2 * cd = reinterpret_cast<double2 *>(smem + L::CELL_OFF);
  double * gxw = reinterpret_cast<double *>(smem + L::GEO_OFF);
  double * gyw = gxw + B;
  double * gzw = gyw + B;
  const uint32_t * pt = reinterpret_cast<const uint32_t *>(smem + L::PT_OFF);
  const int32_t * po = reinterpret_cast<const int32_t *>(smem + L::PO_OFF);

  const int oct = ds / a.nchunks;
  const int chunk = ds - oct * a.nchunks;
  // Octant bits (TransportBase.h:22-32): 3D 4=-mu 2=-eta 1=-xi; 2D quadrant
  // slots PPP, PMP, MPP, MMP: 2=-mu 1=-eta.
  const bool negx = DIM == 3 ? (oct & 4) : (oct & 2);
  const bool negy = DIM == 3 ? (oct & 2) : (oct & 1);
  const bool negz = DIM == 3 ? (oct & 1) : false;

  const int tid = threadIdx.x;
  const int dd = tid % DL;
  const int slot = tid / DL;
  const int n = chunk * a.chunk + dd;
  const bool valid = dd < a.chunk && n < a.dirs;
  double w = 0.0;
  R amu = R(1), aeta = R(1), axi = R(1);
  if (valid)
  {
    const double * t = a.dirtab + 4 * (oct * a.dirs + n);
    amu = R(t[0]);
    aeta = R(t[1]);
    axi = R(t[2]);
    w = t[3];
  }

  const int i0 = bi * B, j0 = bj * B, k0 = bk * B;
  R * gx = reinterpret_cast<R *>(a.fx);
  R * gy = reinterpret_cast<R *>(a.fy);
  R * gz = reinterpret_cast<R *>(a.fz);
  size_t xbase, ybase, zbase = 0;
  if (DIM == 3)
  {
    xbase = (size_t)ds * a.nyp * a.nzp;
    ybase = (size_t)ds * a.nxp * a.nzp;
    zbase = (size_t)ds * a.nxp * a.nyp;
  }
  else
  {
    xbase = (size_t)ds * a.nyp;
    ybase = (size_t)ds * a.nxp;
  }

  // The previous task of this work-group may still be reading LDS.
  __syncthreads();

  // Inflow faces: the upwind brick's outflow, or vacuum (zero) on the boundary.
  for (int e = tid; e < FS * DL; e += NT)
  {
    const int s = e / DL, l = e - s * DL;
    const int u = DIM == 3 ? s / B : s;
    const int v = DIM == 3 ? s % B : 0;
    if (DIM == 3)
    {
      sx[e] = bi > 0 ? gx[(xbase + (size_t)(j0 + u) * a.nzp + (k0 + v)) * DL + l] : R(0);
      sy[e] = bj > 0 ? gy[(ybase + (size_t)(i0 + u) * a.nzp + (k0 + v)) * DL + l] : R(0);
      sz[e] = bk > 0 ? gz[(zbase + (size_t)(i0 + u) * a.nyp + (j0 + v)) * DL + l] : R(0);
    }
    else
    {
      sx[e] = bi > 0 ? gx[(xbase + (size_t)(j0 + u)) * DL + l] : R(0);
      sy[e] = bj > 0 ? gy[(ybase + (size_t)(i0 + u)) * DL + l] : R(0);
    }
  }
  // Cell sources and cross sections, brick-local order (lk*B + lj)*B + li.
  for (int e = tid; e < L::NC; e += NT)
  {
    const int li = e % B;
    const int lj = (e / B) % B;
    const int lk = DIM == 3 ? e / (B * B) : 0;
    const int oi = i0 + li, oj = j0 + lj, ok = k0 + lk;
    double2 v = make_double2(0.0, 1.0);
    if (oi < a.nx && oj < a.ny && ok < a.nz)
    {
      const int i = negx ? a.nx - 1 - oi : oi;
      const int j = negy ? a.ny - 1 - oj : oj;
      const int k = negz ? a.nz - 1 - ok : ok;
      const int c = (k * a.ny + j) * a.nx + i;
      v = make_double2(a.q[c], a.sig[c]);
    }
    cd[e] = v;
  }
  for (int e = tid; e < 3 * B; e += NT)
  {
    const int ax = e / B, l = e - ax * B;
    double v = 1.0;
    if (ax == 0 && i0 + l < a.nx)
      v = a.lx[negx ? a.nx - 1 - (i0 + l) : i0 + l];
    else if (ax == 1 && j0 + l < a.ny)
      v = a.ly[negy ? a.ny - 1 - (j0 + l) : j0 + l];
    else if (ax == 2 && DIM == 3 && k0 + l < a.nz)
      v = a.lz[negz ? a.nz - 1 - (k0 + l) : k0 + l];
    gxw[e] = v;
  }
  __syncthreads();

  for (int p = 0; p < L::NPLANES; ++p)
  {
    const int pe = po[p + 1];
    for (int pos = po[p] + slot; pos < pe; pos += CPP)
    {
      const uint32_t pc = pt[pos];
      const int li = pc & 0xff, lj = (pc >> 8) & 0xff, lk = (pc >> 16) & 0xff;
      const int oi = i0 + li, oj = j0 + lj, ok = k0 + lk;
      if (oi >= a.nx || oj >= a.ny || ok >= a.nz)
        continue;  // outside a partial brick: the whole lane group skips
      const double2 cv = cd[(lk * B + lj) * B + li];
      const double lxd = gxw[li], lyd = gyw[lj];
      R psi;
      if (DIM == 3)
      {
        const double lzd = gzw[lk];
        R & xf = sx[(lj * B + lk) * DL + dd];
        R & yf = sy[(li * B + lk) * DL + dd];
        R & zf = sz[(li * B + lj) * DL + dd];
        R x = xf, y = yf, z = zf;
        if (SCHEME == DOMA_SCHEME_TW_DD)
        {
          if (EXACT)
            psi = twdd3d_exact<R>(R(cv.x), R(cv.y), R(lxd), R(lyd), R(lzd), amu, aeta, axi, x, y, z);
          else
          {
            const double V = lxd * lyd * lzd;
            psi = twdd3d_fast<R>(R(cv.x * V), R(cv.y * V), amu * R(lyd * lzd), aeta * R(lxd * lzd),
                                 axi * R(lxd * lyd), x, y, z);
          }
        }
        else
        {
          if (EXACT)
            psi = dd3d_exact<R>(R(cv.x), R(cv.y), R(lxd), R(lyd), R(lzd), amu, aeta, axi, x, y, z);
          else
            psi = dd3d_fast<R>(R(cv.x), R(cv.y), R(2) * amu * R(1.0 / lxd), R(2) * aeta * R(1.0 / lyd),
                               R(2) * axi * R(1.0 / lzd), x, y, z);
        }
        xf = x;
        yf = y;
        zf = z;
      }
      else
      {
        R & xf = sx[lj * DL + dd];
        R & yf = sy[li * DL + dd];
        R x = xf, y = yf;
        if (SCHEME == DOMA_SCHEME_TW_DD)
        {
          if (EXACT)
            psi = twdd2d_exact<R>(R(cv.x), R(cv.y), R(lxd), R(lyd), amu, aeta, axi, x, y);
          else
          {
            const double V = lxd * lyd;
            psi = twdd2d_fast<R>(R(cv.x * V), R(cv.y * V), amu * R(lyd), aeta * R(lxd), axi, x, y);
          }
        }
        else
        {
          if (EXACT)
            psi = dd2d_exact<R>(R(cv.x), R(cv.y), R(lxd), R(lyd), amu, aeta, x, y);
          else
          {
            R z = R(0);
            psi = dd3d_fast<R>(R(cv.x), R(cv.y), R(2) * amu * R(1.0 / lxd), R(2) * aeta * R(1.0 / lyd),
                               R(0), x, y, z);
          }
        }
        xf = x;
        yf = y;
      }
      double s = valid ? w * static_cast<double>(psi) : 0.0;
      s = group_sum<DL>(s);
      if (dd == 0)
      {
        const int i = negx ? a.nx - 1 - oi : oi;
        const int j = negy ? a.ny - 1 - oj : oj;
        const int k = negz ? a.nz - 1 - ok : ok;
        a.part[(size_t)ds * a.N + (k * a.ny + j) * a.nx + i] = s;
      }
    }
    __syncthreads();
  }

  // Outflow faces for the downwind bricks (padded extents keep every index in range).
  for (int e = tid; e < FS * DL; e += NT)
  {
    const int s = e / DL, l = e - s * DL;
    const int u = DIM == 3 ? s / B : s;
    const int v = DIM == 3 ? s % B : 0;
    if (DIM == 3)
    {
      gx[(xbase + (size_t)(j0 + u) * a.nzp + (k0 + v)) * DL + l] = sx[e];
      gy[(ybase + (size_t)(i0 + u) * a.nzp + (k0 + v)) * DL + l] = sy[e];
      gz[(zbase + (size_t)(i0 + u) * a.nyp + (j0 + v)) * DL + l] = sz[e];
    }
    else
    {
      gx[(xbase + (size_t)(j0 + u)) * DL + l] = sx[e];
      gy[(ybase + (size_t)(i0 + u)) * DL + l] = sy[e];
    }
  }
}

// One launch per brick diagonal.
template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
__global__ __launch_bounds__(NT) void
brick_sweep_kernel(BrickArgs a)
{
  extern __shared__ __align__(16) unsigned char smem[];
  load_plane_table<R, DIM, B, DL, NT>(a, smem);
  const int ds = blockIdx.x % a.ndirsets;
  const int br = a.diag_begin + blockIdx.x / a.ndirsets;
  sweep_brick<R, DIM, SCHEME, EXACT, B, DL, NT>(a, a.bricks[3 * br + 0], a.bricks[3 * br + 1],
                                               a.bricks[3 * br + 2], ds, smem);
}

__device__ __forceinline__ int
brick_linear(const BrickArgs & a, int bi, int bj, int bk)
{
  return (bk * a.nty + bj) * a.ntx + bi;
}

// Bounded wait for an upwind brick's epoch flag (thread 0 only).  Polls with
// relaxed device-coherent loads (the caller issues one acquire fence after all
// waits); gives up at a wall-clock deadline, or at once when another wait of
// this sweep has already expired, so a broken schedule ends the launch quickly.
__device__ __forceinline__ void
wait_flag(const BrickArgs & a, const uint32_t * f)
{
  if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= a.epoch)
    return;
  const long long t0 = wall_clock64();
  int polls = 0;
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < a.epoch)
  {
    __builtin_amdgcn_s_sleep(2);
    if ((++polls & 31) == 0)
    {
      if (__hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        return;
      if (wall_clock64() - t0 > a.wait_ticks)
      {
        atomicExch(a.error, 1);
        return;
      }
    }
  }
}

// One persistent launch per sweep; tasks in ticket (= diagonal) order.
template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
__global__ __launch_bounds__(NT) void
brick_flow_kernel(BrickArgs a)
{
  using L = Layout<R, DIM, B, DL>;
  extern __shared__ __align__(16) unsigned char smem[];
  int * s_task = reinterpret_cast<int *>(smem + L::TASK_OFF);
  load_plane_table<R, DIM, B, DL, NT>(a, smem);
  const int nbricks = a.ntx * a.nty * a.ntz;
  while (true)
  {
    // A barrier heads every iteration so that the divergent "publish flag" of
    // one task and "take ticket" of the next can never be merged across the
    // back-edge into a nested loop that lane 0 leaves without its wave.
    __syncthreads();
    if (threadIdx.x == 0)
    {
      const unsigned long long t = atomicAdd(a.ticket, 1ull) - a.ticket_base;
      *s_task = t < (unsigned long long)a.ntasks ? (int)t : -1;
    }
    __syncthreads();
    // The task index must be provably uniform: a loop exit that the compiler
    // sees as divergent is lowered with per-lane exec masks, which would let
    // lane 0 leave for the next ticket while the rest of its wave loops back to
    // the barrier.
    const int t = __builtin_amdgcn_readfirstlane(*s_task);
    if (t < 0)
      break;
    const int br = t / a.ndirsets;
    const int ds = t - br * a.ndirsets;
    const int bi = a.bricks[3 * br + 0], bj = a.bricks[3 * br + 1], bk = a.bricks[3 * br + 2];
    uint32_t * flags = a.flags + (size_t)ds * nbricks;
    if (threadIdx.x == 0)
    {
      if (bi > 0)
        wait_flag(a, flags + brick_linear(a, bi - 1, bj, bk));
      if (bj > 0)
        wait_flag(a, flags + brick_linear(a, bi, bj - 1, bk));
      if (bk > 0)
        wait_flag(a, flags + brick_linear(a, bi, bj, bk - 1));
    }
    __syncthreads();
    __threadfence();
    sweep_brick<R, DIM, SCHEME, EXACT, B, DL, NT>(a, bi, bj, bk, ds, smem);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(flags + brick_linear(a, bi, bj, bk), a.epoch, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------------------- //
// 1D sweep: one thread per direction, cells in sweep order                   //
// (TransportSolver1D.cpp:840-915).                                           //
// ------------------------------------------------------------------------- //

template <typename R>
__global__ __launch_bounds__(1024) void
sweep1d_kernel(Sweep1DArgs a)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * a.dirs)
    return;
  const int half = t / a.dirs;
  const double * d = a.dirtab + 4 * t;
  const R amu = R(d[0]);
  const double w = d[3];
  R f = R(0);
  for (int ii = 0; ii < a.nx; ++ii)
  {
    const int i = half == 0 ? ii : a.nx - 1 - ii;
    const R psi = dd1d<R>(R(a.q[i]), R(a.sig[i]), R(a.lx[i]), amu, f);
    a.part[(size_t)t * a.nx + i] = w * static_cast<double>(psi);
  }
}

// ------------------------------------------------------------------------- //
// Launch plumbing                                                            //
// ------------------------------------------------------------------------- //

template <typename K>
int
prepare(K k, size_t shmem, bool & done)
{
  if (done)
    return DOMA_OK;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  if (e != hipSuccess)
    return hip_fail(e, "hipFuncSetAttribute(sweep kernel)");
  done = true;
  return DOMA_OK;
}

template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
int
launch_one(const BrickArgs & a, int schedule, int grid_flow, hipStream_t s)
{
  using L = Layout<R, DIM, B, DL>;
  const size_t shmem = L::BYTES;
  if (schedule == 0)
  {
    auto k = brick_sweep_kernel<R, DIM, SCHEME, EXACT, B, DL, NT>;
    static bool ready = false;
    int rc = prepare(k, shmem, ready);
    if (rc)
      return rc;
    const unsigned grid = (unsigned)a.diag_count * (unsigned)a.ndirsets;
    hipLaunchKernelGGL(k, dim3(grid), dim3(NT), shmem, s, a);
  }
  else
  {
    auto k = brick_flow_kernel<R, DIM, SCHEME, EXACT, B, DL, NT>;
    static bool ready = false;
    int rc = prepare(k, shmem, ready);
    if (rc)
      return rc;
    hipLaunchKernelGGL(k, dim3(grid_flow), dim3(NT), shmem, s, a);
  }
  return DOMA_OK;
}

template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
int
occupancy_one(int * per_cu)
{
  using L = Layout<R, DIM, B, DL>;
  auto k = brick_flow_kernel<R, DIM, SCHEME, EXACT, B, DL, NT>;
  static bool ready = false;
  int rc = prepare(k, L::BYTES, ready);
  if (rc)
    return rc;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, NT, L::BYTES);
  if (e != hipSuccess)
    return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  return DOMA_OK;
}

// Calls F::template run<R, DIM, SCHEME, EXACT, B, DL, NT>() for the requested shape.
template <typename F, typename R, int DIM, int SCHEME, bool EXACT, int B, int NT>
int
by_dl(F & f, int dl)
{
  switch (dl)
  {
    case 8: return f.template run<R, DIM, SCHEME, EXACT, B, 8, NT>();
    case 16: return f.template run<R, DIM, SCHEME, EXACT, B, 16, NT>();
    case 32: return f.template run<R, DIM, SCHEME, EXACT, B, 32, NT>();
    case 64: return f.template run<R, DIM, SCHEME, EXACT, B, 64, NT>();
  }
  return fail(DOMA_ERR_INVALID, "unsupported direction-lane width");
}

template <typename F, typename R, int DIM, int SCHEME, bool EXACT>
int
by_shape(F & f, int brick, int dl, int threads)
{
  if (DIM == 3)
  {
    if (brick == 4 && threads == 128) return by_dl<F, R, DIM, SCHEME, EXACT, 4, 128>(f, dl);
    if (brick == 4 && threads == 256) return by_dl<F, R, DIM, SCHEME, EXACT, 4, 256>(f, dl);
    if (brick == 8 && threads == 256) return by_dl<F, R, DIM, SCHEME, EXACT, 8, 256>(f, dl);
    if (brick == 8 && threads == 512) return by_dl<F, R, DIM, SCHEME, EXACT, 8, 512>(f, dl);
  }
  else
  {
    if (brick == 8 && threads == 256) return by_dl<F, R, DIM, SCHEME, EXACT, 8, 256>(f, dl);
    if (brick == 16 && threads == 256) return by_dl<F, R, DIM, SCHEME, EXACT, 16, 256>(f, dl);
  }
  return fail(DOMA_ERR_INVALID, "unsupported brick / work-group shape");
}

template <typename F>
int
dispatch(F & f, int dims, int scheme, int precision, int exact, int brick, int dl, int threads)
{
  const bool tw = scheme == DOMA_SCHEME_TW_DD;
  const bool f32 = precision == DOMA_FP32_SWEEP;
#define DOMA_DISPATCH(R, D)                                                                        \
  return tw ? (exact ? by_shape<F, R, D, DOMA_SCHEME_TW_DD, true>(f, brick, dl, threads)           \
                     : by_shape<F, R, D, DOMA_SCHEME_TW_DD, false>(f, brick, dl, threads))         \
            : (exact ? by_shape<F, R, D, DOMA_SCHEME_DD, true>(f, brick, dl, threads)              \
                     : by_shape<F, R, D, DOMA_SCHEME_DD, false>(f, brick, dl, threads))
  if (dims == 3)
  {
    if (f32)
      DOMA_DISPATCH(float, 3);
    DOMA_DISPATCH(double, 3);
  }
  if (f32)
    DOMA_DISPATCH(float, 2);
  DOMA_DISPATCH(double, 2);
#undef DOMA_DISPATCH
}

struct LaunchF
{
  const BrickArgs & a;
  int schedule;
  int grid_flow;
  hipStream_t s;
  template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
  int run()
  {
    return launch_one<R, DIM, SCHEME, EXACT, B, DL, NT>(a, schedule, grid_flow, s);
  }
};

struct OccF
{
  int * per_cu;
  template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
  int run()
  {
    return occupancy_one<R, DIM, SCHEME, EXACT, B, DL, NT>(per_cu);
  }
};
} // namespace

bool
brick_variant_exists(int dims, int brick, int dl, int threads)
{
  if (dl != 8 && dl != 16 && dl != 32 && dl != 64)
    return false;
  if (threads < dl)
    return false;
  if (dims == 3)
    return (brick == 4 && (threads == 256 || threads == 128)) ||
           (brick == 8 && (threads == 256 || threads == 512));
  if (dims == 2)
    return (brick == 8 || brick == 16) && threads == 256;
  return false;
}

int
launch_brick_sweep(const BrickArgs & a, int dims, int scheme, int precision, int exact, int brick,
                   int dl, int threads, int schedule, int grid_flow, hipStream_t s)
{
  if (!brick_variant_exists(dims, brick, dl, threads))
    return fail(DOMA_ERR_INVALID, "no sweep kernel for this brick / lane / work-group shape");
  LaunchF f{a, schedule, grid_flow, s};
  return dispatch(f, dims, scheme, precision, exact, brick, dl, threads);
}

int
flow_blocks_per_cu(int dims, int scheme, int precision, int exact, int brick, int dl, int threads,
                   int * per_cu)
{
  if (!brick_variant_exists(dims, brick, dl, threads))
    return fail(DOMA_ERR_INVALID, "no sweep kernel for this brick / lane / work-group shape");
  OccF f{per_cu};
  return dispatch(f, dims, scheme, precision, exact, brick, dl, threads);
}

int
launch_sweep1d(const Sweep1DArgs & a, int precision, hipStream_t s)
{
  const int nthreads = 2 * a.dirs;
  const int block = nthreads <= 1024 ? ((nthreads + 63) / 64) * 64 : 1024;
  const int grid = (nthreads + block - 1) / block;
  if (precision == DOMA_FP32_SWEEP)
    hipLaunchKernelGGL(sweep1d_kernel<float>, dim3(grid), dim3(block), 0, s, a);
  else
    hipLaunchKernelGGL(sweep1d_kernel<double>, dim3(grid), dim3(block), 0, s, a);
  return DOMA_OK;
}
} // namespace doma
