// sweep.hip -- launch plumbing of the brick-wavefront sweep (kernel bodies in
// sweep_impl.h, instantiated per precision and scheme by sweep_{f64,f32}_{tw,dd}.hip)
// and the 1D sweep kernel.
#include <hip/hip_runtime.h>

#include "sweep_impl.h"

namespace doma
{
namespace sweepk
{
template <>
int dispatch_shapes<double, DOMA_SCHEME_TW_DD>(const BrickArgs &, int, int, int, int, int,
                                              hipStream_t, int *);
template <>
int dispatch_shapes<double, DOMA_SCHEME_DD>(const BrickArgs &, int, int, int, int, int,
                                           hipStream_t, int *);
template <>
int dispatch_shapes<float, DOMA_SCHEME_TW_DD>(const BrickArgs &, int, int, int, int, int,
                                             hipStream_t, int *);
template <>
int dispatch_shapes<float, DOMA_SCHEME_DD>(const BrickArgs &, int, int, int, int, int,
                                          hipStream_t, int *);
template <>
int dispatch_batch_shapes<double, DOMA_SCHEME_TW_DD>(const BatchArgs &, int, int, int,
                                                    hipStream_t, int *);
template <>
int dispatch_batch_shapes<double, DOMA_SCHEME_DD>(const BatchArgs &, int, int, int, hipStream_t,
                                                 int *);
template <>
int dispatch_batch_shapes<float, DOMA_SCHEME_TW_DD>(const BatchArgs &, int, int, int,
                                                   hipStream_t, int *);
template <>
int dispatch_batch_shapes<float, DOMA_SCHEME_DD>(const BatchArgs &, int, int, int, hipStream_t,
                                                int *);
} // namespace sweepk

namespace
{
int
dispatch_batch(const BatchArgs & b, int scheme, int precision, int shape, int op, int grid,
               hipStream_t s, int * per_cu)
{
  using namespace sweepk;
  const bool f32 = precision == DOMA_FP32_SWEEP;
  if (scheme == DOMA_SCHEME_TW_DD)
    return f32 ? dispatch_batch_shapes<float, DOMA_SCHEME_TW_DD>(b, shape, op, grid, s, per_cu)
               : dispatch_batch_shapes<double, DOMA_SCHEME_TW_DD>(b, shape, op, grid, s, per_cu);
  return f32 ? dispatch_batch_shapes<float, DOMA_SCHEME_DD>(b, shape, op, grid, s, per_cu)
             : dispatch_batch_shapes<double, DOMA_SCHEME_DD>(b, shape, op, grid, s, per_cu);
}

int
dispatch(const BrickArgs & a, int scheme, int precision, int exact, int shape, int op,
         int schedule, int grid, hipStream_t s, int * per_cu)
{
  using namespace sweepk;
  const bool f32 = precision == DOMA_FP32_SWEEP;
  if (scheme == DOMA_SCHEME_TW_DD)
    return f32 ? dispatch_shapes<float, DOMA_SCHEME_TW_DD>(a, exact, shape, op, schedule, grid, s,
                                                           per_cu)
               : dispatch_shapes<double, DOMA_SCHEME_TW_DD>(a, exact, shape, op, schedule, grid,
                                                            s, per_cu);
  return f32 ? dispatch_shapes<float, DOMA_SCHEME_DD>(a, exact, shape, op, schedule, grid, s,
                                                      per_cu)
             : dispatch_shapes<double, DOMA_SCHEME_DD>(a, exact, shape, op, schedule, grid, s,
                                                       per_cu);
}

// 1D sweep: one thread per direction, cells in sweep order
// (TransportSolver1D.cpp:840-915).  With STAGE the slab's source, cross section
// and widths are first copied to LDS, so each cell step reads LDS instead of
// waiting on global memory.
constexpr int kStage1D = 2048;  // cells staged in LDS (3 x 8 B each)

template <typename R, bool STAGE>
__global__ __launch_bounds__(1024) void
sweep1d_kernel(Sweep1DArgs a)
{
  if (a.skip && *a.skip)
    return;
  __shared__ double s_buf[STAGE ? 3 * kStage1D : 1];
  const double * q = a.q;
  const double * sg = a.sig;
  const double * lx = a.lx;
  if constexpr (STAGE)
  {
    for (int i = threadIdx.x; i < a.nx; i += blockDim.x)
    {
      s_buf[i] = a.q[i];
      s_buf[kStage1D + i] = a.sig[i];
      s_buf[2 * kStage1D + i] = a.lx[i];
    }
    __syncthreads();
    q = s_buf;
    sg = s_buf + kStage1D;
    lx = s_buf + 2 * kStage1D;
  }
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
    const R psi = dd1d<R>(R(q[i]), R(sg[i]), R(lx[i]), amu, f);
    a.part[(size_t)t * a.nx + i] = w * static_cast<double>(psi);
  }
}
} // namespace

bool
shape_info(int shape, ShapeInfo * out)
{
  if (shape < 0 || shape >= sweepk::kNumShapes)
    return false;
  const auto & s = sweepk::kShapes[shape];
  out->dims = s.dims;
  out->brick = s.B;
  out->dl = s.DL;
  out->ndl = s.NDL;
  out->ko = s.KO;
  out->threads = s.NT;
  return true;
}

int
num_shapes()
{
  return sweepk::kNumShapes;
}

size_t
shape_lds_bytes(int shape, int precision)
{
  ShapeInfo si;
  if (!shape_info(shape, &si))
    return 0;
  const size_t rs = precision == DOMA_FP32_SWEEP ? 4 : 8;
  const int fs = si.dims == 3 ? si.brick * si.brick : si.brick;
  const int nc = si.dims == 3 ? si.brick * si.brick * si.brick : si.brick * si.brick;
  const int nax = si.dims == 3 ? 3 : 2;
  const int nplanes = si.dims == 3 ? 3 * si.brick - 2 : 2 * si.brick - 1;
  // faces, cell records (48 B in FP64, 24 B in FP32), plane table, plane
  // offsets, inverse plane table, task slot -- the layout of sweep_impl.h
  const size_t faces = rs * fs * si.dl * si.ndl * nax * si.ko;
  const size_t rec = precision == DOMA_FP32_SWEEP ? sizeof(sweepk::CellRec<float>)
                                                  : sizeof(sweepk::CellRec<double>);
  const size_t ptb = sweepk::wide_plane_table(faces, nc, nplanes, si.ko, rec) ? 8 : 4;
  return sweepk::layout_bytes(faces, nc, nplanes, si.ko, ptb, rec);
}

int
launch_brick_sweep(const BrickArgs & a, int scheme, int precision, int exact, int shape,
                   int schedule, int grid, hipStream_t s)
{
  return dispatch(a, scheme, precision, exact, shape, 0, schedule, grid, s, nullptr);
}

int
sweep_blocks_per_cu(int scheme, int precision, int exact, int shape, int schedule, int * per_cu)
{
  BrickArgs dummy{};
  return dispatch(dummy, scheme, precision, exact, shape, 1, schedule, 0, nullptr, per_cu);
}

int
launch_brick_batch(const BatchArgs & b, int scheme, int precision, int shape, int grid,
                   hipStream_t s)
{
  return dispatch_batch(b, scheme, precision, shape, 0, grid, s, nullptr);
}

int
batch_blocks_per_cu(int scheme, int precision, int shape, int * per_cu)
{
  BatchArgs dummy{};
  return dispatch_batch(dummy, scheme, precision, shape, 1, 0, nullptr, per_cu);
}

int
launch_sweep1d(const Sweep1DArgs & a, int precision, hipStream_t s)
{
  const int nthreads = 2 * a.dirs;
  const int block = nthreads <= 1024 ? ((nthreads + 63) / 64) * 64 : 1024;
  const int grid = (nthreads + block - 1) / block;
  const bool stage = a.nx <= kStage1D;
  if (precision == DOMA_FP32_SWEEP)
  {
    if (stage)
      hipLaunchKernelGGL((sweep1d_kernel<float, true>), dim3(grid), dim3(block), 0, s, a);
    else
      hipLaunchKernelGGL((sweep1d_kernel<float, false>), dim3(grid), dim3(block), 0, s, a);
  }
  else if (stage)
    hipLaunchKernelGGL((sweep1d_kernel<double, true>), dim3(grid), dim3(block), 0, s, a);
  else
    hipLaunchKernelGGL((sweep1d_kernel<double, false>), dim3(grid), dim3(block), 0, s, a);
  return DOMA_OK;
}
} // namespace doma
