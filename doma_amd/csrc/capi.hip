// capi.hip -- the extern "C" entry points declared in include/doma_c.h: context
// and device memory management, the brick schedule of the sweep, and the source
// iteration / multigroup Gauss-Seidel drivers (TransportSolver3D.cpp:262-563 and
// the 2D/1D counterparts) with all per-cell state resident in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "internal.h"

namespace
{
thread_local std::string g_last_error;

template <typename T>
int
dev_alloc(T ** p, size_t count)
{
  *p = nullptr;
  if (count == 0)
    count = 1;
  DOMA_HIP(hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count));
  return DOMA_OK;
}

template <typename T>
void
dev_free(T *& p)
{
  if (p)
    (void)hipFree(reinterpret_cast<void *>(p));
  p = nullptr;
}

int
next_pow2(int v)
{
  int p = 1;
  while (p < v)
    p <<= 1;
  return p;
}

// Autotuner cache: (dims, n, G-independent shape, dirs, precision, scheme) -> tuning.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int, int, int>, doma_tuning> g_tune_cache;

std::tuple<int, int, int, int, int, int, int>
tune_key(const doma_ctx * c)
{
  return std::make_tuple(c->dims, c->nx, c->ny, c->nz, c->dirs, c->precision, c->scheme);
}
} // namespace

namespace doma
{
int
fail(int code, const std::string & msg)
{
  g_last_error = msg;
  return code;
}

int
hip_fail(hipError_t e, const char * what)
{
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return DOMA_ERR_DEVICE;
}

IterArgs
iter_args(doma_ctx * c, int g)
{
  IterArgs a{};
  a.dims = c->dims;
  a.N = c->N;
  a.nx = c->nx;
  a.ny = c->ny;
  a.G = c->G;
  a.g = g;
  a.ndirsets = c->ndirsets;
  a.mat = c->d_mat;
  a.sigt = c->d_sigt;
  a.sigs = c->d_sigs;
  a.src = c->d_src;
  a.has_src = c->d_has_src;
  a.lx = c->d_lx;
  a.ly = c->d_ly;
  a.lz = c->d_lz;
  a.sig = c->d_sig;
  a.q = c->d_q;
  a.qmg = c->d_qmg;
  a.phi = c->d_phi;
  a.prev = c->d_prev;
  a.part = c->d_part;
  a.dphi = c->d_dphi;
  a.red = c->d_red;
  a.res = c->d_res;
  a.err = c->d_err;
  if (c->transient)
  {
    a.inv_v = c->d_inv_v;
    a.nd = c->d_nd;
    a.chi_d = c->d_chi_d;
    a.beta = c->d_beta;
    a.lambda = c->d_lambda;
    a.decay = c->d_decay;
    if (c->steps_set)
    {
      a.step_type = c->d_step_type;
      a.step_in = c->d_step_in;
      a.step_out = c->d_step_out;
      a.step_src = c->d_step_src;
    }
    a.last = c->d_last;
    a.dnp = c->d_dnp;
    a.t = c->tr_t;
    a.dt = c->tr_dt;
  }
  if (c->fission)
  {
    a.chi = c->d_chi;
    a.prod = c->d_prod;
    a.prompt = c->d_prompt;
    a.has_chi = c->d_has_chi;
    a.has_prod = c->d_has_prod;
  }
  return a;
}

void
free_sweep(doma_ctx * c)
{
  dev_free(c->d_bricks);
  dev_free(c->d_plane_cells);
  dev_free(c->d_plane_off);
  dev_free(c->d_fx);
  dev_free(c->d_fy);
  dev_free(c->d_fz);
  dev_free(c->d_part);
  dev_free(c->d_flags);
  dev_free(c->d_pencils);
  c->npencils = 0;
}

// Reports (and clears) an expired dependency wait of the dataflow schedule.
int
check_sweep_error(doma_ctx * c)
{
  int h = 0;
  DOMA_HIP(hipMemcpyAsync(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DOMA_HIP(hipStreamSynchronize(c->stream));
  if (h == 0)
    return DOMA_OK;
  return reset_after_sweep_error(c);
}

int
reset_after_sweep_error(doma_ctx * c)
{
  DOMA_HIP(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
  DOMA_HIP(hipMemsetAsync(c->d_ticket, 0,
                          sizeof(unsigned long long) * kTicketQueues * kTicketStride, c->stream));
  if (c->d_flags)
    DOMA_HIP(hipMemsetAsync(c->d_flags, 0, sizeof(uint32_t) * (size_t)c->ntasks, c->stream));
  DOMA_HIP(hipStreamSynchronize(c->stream));
  c->ticket_base = 0;
  c->epoch = 0;
  return fail(DOMA_ERR_DEVICE, "sweep dataflow wait expired: an upwind brick never finished");
}

// Direction sets a shape implies: at most min(64, DL * NDL) directions per set,
// spread evenly.
void
shape_chunking(int shape, int dirs, int * chunk, int * nchunks)
{
  ShapeInfo si{};
  shape_info(shape, &si);
  const int w = std::max(1, std::min(64, si.dl * si.ndl));
  *nchunks = (dirs + w - 1) / w;
  *chunk = (dirs + *nchunks - 1) / *nchunks;
}

// Kernel shape for a problem: shapes of the right dimension whose octant group
// divides the octant count and that agree with every non-zero tuning hint.
// Without hints the default prefers the fewest direction sets (one per 64
// directions) and the narrowest direction width that covers them, then 8^3
// bricks on large grids (4^3 otherwise; 16^2 / 8^2 tiles in 2D), then one
// direction per lane and one octant per group (measured fastest on MI355X: the
// sweep is latency bound, and those shapes give the shortest per-brick critical
// path at the highest occupancy).  Shapes that split the directions further are
// left to the autotuner.
int
select_shape(int dims, int noct, int dirs, int minext, const doma_tuning & t)
{
  if (t.shape > 0)
  {
    ShapeInfo si{};
    if (!shape_info(t.shape - 1, &si) || si.dims != dims || noct % si.ko != 0)
      return -1;
    return t.shape - 1;
  }
  const int pref_brick = dims == 3 ? (minext >= 96 ? 8 : 4) : (minext >= 128 ? 16 : 8);
  const int min_sets = (dirs + 63) / 64;
  int best = -1;
  long best_score = 0;
  for (int s = 0; s < num_shapes(); ++s)
  {
    ShapeInfo si{};
    shape_info(s, &si);
    if (si.dims != dims || noct % si.ko != 0)
      continue;
    if ((t.brick && si.brick != t.brick) || (t.threads && si.threads != t.threads) ||
        (t.octants && si.ko != t.octants) || (t.dirs_per_lane && si.ndl != t.dirs_per_lane))
      continue;
    int chunk = 0, nchunks = 0;
    shape_chunking(s, dirs, &chunk, &nchunks);
    const int dw = si.dl * si.ndl;
    const int pref_threads = si.brick == 8 ? 512 : 256;
    const long score = (long)(nchunks - min_sets) * 100000 + (long)(dw - chunk) * 1000 +
                       (si.brick != pref_brick ? 100 : 0) + (si.ndl == 1 ? 0 : 20) +
                       (si.ko - 1) * 4 + (si.threads != pref_threads ? 1 : 0);
    if (best < 0 || score < best_score)
    {
      best = s;
      best_score = score;
    }
  }
  return best;
}

// Brick schedule, hyperplane tables and face/partial buffers for the current
// tuning.  1D uses one thread per direction and no bricks.
int
configure_sweep(doma_ctx * c)
{
  free_sweep(c);
  if (c->dims == 1)
  {
    c->ndirsets = 2 * c->dirs;
    c->launches_per_sweep = 1;
    return dev_alloc(&c->d_part, (size_t)c->ndirsets * c->N);
  }
  // 2D: one lane per row (schedule 2) when asked for, or by default when no
  // brick shape or schedule is requested and the mesh fits one work-group.
  const bool untuned = !c->tuning.shape && !c->tuning.schedule && !c->tuning.brick &&
                       !c->tuning.threads && !c->tuning.octants && !c->tuning.dirs_per_lane;
  if (c->dims == 2 && (c->tuning.schedule == 2 || (untuned && rows2d_fits(c->nx, c->ny))))
  {
    if (!rows2d_fits(c->nx, c->ny))
      return fail(DOMA_ERR_INVALID, "the row-pipelined 2D sweep needs ny <= " +
                                        std::to_string(kRowsMaxThreads));
    c->shape = -1;
    c->brick = 0;
    c->threads = ((c->ny + 63) / 64) * 64;
    c->dl = c->ndl = c->ko = c->dw = 1;
    c->chunk = 1;
    c->nchunks = c->dirs;
    c->exact = c->tuning.exact_arith ? 1 : 0;
    c->schedule = 2;
    c->ndirsets = c->noct * c->dirs;
    c->launches_per_sweep = 1;
    c->ntasks = 0;
    return dev_alloc(&c->d_part, (size_t)c->ndirsets * c->N);
  }
  if (c->dims == 3 && c->tuning.schedule == 2)
    return fail(DOMA_ERR_INVALID, "schedule 2 (one lane per row) is a 2D sweep");
  if (c->tuning.schedule == 3 && c->dims != 3)
    return fail(DOMA_ERR_INVALID, "schedule 3 (pencils of bricks along z) is a 3D sweep");
  if (c->tuning.schedule < 0 || c->tuning.schedule > 3)
    return fail(DOMA_ERR_INVALID, "unknown sweep schedule " + std::to_string(c->tuning.schedule));
  const int minext = c->dims == 3 ? std::min(c->nx, std::min(c->ny, c->nz)) : std::min(c->nx, c->ny);
  const int shape = select_shape(c->dims, c->noct, c->dirs, minext, c->tuning);
  ShapeInfo si{};
  if (shape < 0 || !shape_info(shape, &si))
    return fail(DOMA_ERR_INVALID, "no sweep kernel shape matches the tuning for " +
                                      std::to_string(c->dirs) + " directions per octant");
  shape_chunking(shape, c->dirs, &c->chunk, &c->nchunks);
  const int B = si.brick;
  c->shape = shape;
  c->brick = B;
  c->threads = si.threads;
  c->dl = si.dl;
  c->ndl = si.ndl;
  c->ko = si.ko;
  c->dw = si.dl * si.ndl;
  c->exact = c->tuning.exact_arith ? 1 : 0;
  c->schedule = c->tuning.schedule == 1 || c->tuning.schedule == 3 ? c->tuning.schedule : 0;
  c->ndirsets = c->noct * c->nchunks;

  const int ntx = (c->nx + B - 1) / B, nty = (c->ny + B - 1) / B;
  const int ntz = c->dims == 3 ? (c->nz + B - 1) / B : 1;
  c->nxp = ntx * B;
  c->nyp = nty * B;
  c->nzp = c->dims == 3 ? ntz * B : 1;

  // Bricks sorted by diagonal bi+bj+bk (oriented coordinates).
  const int ndiag = ntx + nty + ntz - 2;
  std::vector<int16_t> bricks;
  c->diag_off.assign(ndiag + 1, 0);
  for (int s = 0; s < ndiag; ++s)
  {
    c->diag_off[s] = (int)(bricks.size() / 3);
    for (int bk = 0; bk < ntz; ++bk)
      for (int bj = 0; bj < nty; ++bj)
      {
        const int bi = s - bj - bk;
        if (bi < 0 || bi >= ntx)
          continue;
        bricks.push_back((int16_t)bi);
        bricks.push_back((int16_t)bj);
        bricks.push_back((int16_t)bk);
      }
  }
  c->diag_off[ndiag] = (int)(bricks.size() / 3);
  c->ntx = ntx;
  c->nty = nty;
  c->ntz = ntz;
  c->ntasks = (int)(bricks.size() / 3) * (c->ndirsets / c->ko);
  c->launches_per_sweep = c->schedule == 0 ? ndiag : 1;

  // Pencils (bi, bj) in diagonal order bi+bj: the upwind pencils of a pencil
  // come before it, as the upwind bricks of a brick do.
  std::vector<int16_t> pencils;
  if (c->schedule == 3)
  {
    for (int s = 0; s < ntx + nty - 1; ++s)
      for (int bj = 0; bj < nty; ++bj)
      {
        const int bi = s - bj;
        if (bi < 0 || bi >= ntx)
          continue;
        pencils.push_back((int16_t)bi);
        pencils.push_back((int16_t)bj);
      }
    c->npencils = (int)(pencils.size() / 2);
  }

  // Brick-local cells ordered by hyperplane li+lj+lk.
  const int bz = c->dims == 3 ? B : 1;
  const int nplanes = c->dims == 3 ? 3 * B - 2 : 2 * B - 1;
  std::vector<uint32_t> cells;
  std::vector<int32_t> off(nplanes + 1, 0);
  for (int p = 0; p < nplanes; ++p)
  {
    off[p] = (int32_t)cells.size();
    for (int lk = 0; lk < bz; ++lk)
      for (int lj = 0; lj < B; ++lj)
      {
        const int li = p - lj - lk;
        if (li < 0 || li >= B)
          continue;
        cells.push_back((uint32_t)((lk * B + lj) * B + li));
      }
  }
  off[nplanes] = (int32_t)cells.size();

  int rc;
  if ((rc = dev_alloc(&c->d_bricks, bricks.size())) ||
      (rc = dev_alloc(&c->d_plane_cells, cells.size())) ||
      (rc = dev_alloc(&c->d_plane_off, off.size())))
    return rc;
  DOMA_HIP(hipMemcpy(c->d_bricks, bricks.data(), bricks.size() * sizeof(int16_t),
                     hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_plane_cells, cells.data(), cells.size() * sizeof(uint32_t),
                     hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_plane_off, off.data(), off.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
  if (c->schedule == 3)
  {
    if ((rc = dev_alloc(&c->d_pencils, pencils.size())))
      return rc;
    DOMA_HIP(hipMemcpy(c->d_pencils, pencils.data(), pencils.size() * sizeof(int16_t),
                       hipMemcpyHostToDevice));
  }

  const size_t rs = c->precision == DOMA_FP32_SWEEP ? sizeof(float) : sizeof(double);
  size_t nfx, nfy, nfz;
  if (c->dims == 3)
  {
    nfx = (size_t)c->ndirsets * c->nyp * c->nzp * c->dw;
    nfy = (size_t)c->ndirsets * c->nxp * c->nzp * c->dw;
    nfz = (size_t)c->ndirsets * c->nxp * c->nyp * c->dw;
  }
  else
  {
    nfx = (size_t)c->ndirsets * c->nyp * c->dw;
    nfy = (size_t)c->ndirsets * c->nxp * c->dw;
    nfz = 1;
  }
  c->face_bytes = (nfx + nfy + nfz) * rs;
  DOMA_HIP(hipMalloc(&c->d_fx, nfx * rs));
  DOMA_HIP(hipMalloc(&c->d_fy, nfy * rs));
  DOMA_HIP(hipMalloc(&c->d_fz, nfz * rs));

  // Dataflow schedule: epoch flags per (direction set, brick), zeroed once; the
  // persistent grid is as many work-groups as fit on the device at once.
  if ((rc = dev_alloc(&c->d_flags, (size_t)c->ntasks)))
    return rc;
  DOMA_HIP(hipMemset(c->d_flags, 0, sizeof(uint32_t) * (size_t)c->ntasks));
  c->epoch = 0;
  // The ticket queues restart with the new task list.
  DOMA_HIP(hipMemset(c->d_ticket, 0, sizeof(unsigned long long) * kTicketQueues * kTicketStride));
  c->ticket_base = 0;
  if (c->schedule == 1 || c->schedule == 3)
  {
    int per_cu = 0;
    if ((rc = sweep_blocks_per_cu(c->scheme, c->precision, c->exact, c->shape, c->schedule,
                                  &per_cu)))
      return rc;
    hipDeviceProp_t prop;
    DOMA_HIP(hipGetDeviceProperties(&prop, c->device));
    // As many queues as divide the direction-set pairs per brick (up to 16), and
    // a resident grid that is a multiple of them.
    const int pairs = c->ndirsets / c->ko;
    c->nqueues = 1;
    for (int nq = kTicketQueues; nq > 1; nq /= 2)
      if (pairs % nq == 0)
      {
        c->nqueues = nq;
        break;
      }
    const int resident = std::max(per_cu, 1) * prop.multiProcessorCount;
    const int tasks = c->schedule == 3 ? c->npencils * pairs : c->ntasks;
    int grid = std::min(tasks, resident);
    grid = std::max(c->nqueues, grid / c->nqueues * c->nqueues);
    if (grid > resident)
    {
      c->nqueues = 1;
      grid = std::max(1, std::min(tasks, resident));
    }
    c->grid_flow = grid;
  }
  return dev_alloc(&c->d_part, (size_t)c->ndirsets * c->N);
}

// Sweep arguments of a 2D/3D context for source q (schedule fields left zero).
BrickArgs
brick_args(doma_ctx * c, const double * q)
{
  BrickArgs a{};
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.N = c->N;
  a.dirs = c->dirs;
  a.chunk = c->chunk;
  a.nchunks = c->nchunks;
  a.ndirsets = c->ndirsets;
  a.dirtab = c->d_dirtab;
  a.lx = c->d_lx;
  a.ly = c->d_ly;
  a.lz = c->d_lz;
  a.rlx = c->d_rlx;
  a.rly = c->d_rly;
  a.rlz = c->d_rlz;
  a.q = q;
  a.sig = c->d_sig;
  a.part = c->d_part;
  a.fx = c->d_fx;
  a.fy = c->d_fy;
  a.fz = c->d_fz;
  a.nxp = c->nxp;
  a.nyp = c->nyp;
  a.nzp = c->nzp;
  a.bricks = c->d_bricks;
  a.plane_cells = c->d_plane_cells;
  a.plane_off = c->d_plane_off;
  a.ntx = c->ntx;
  a.nty = c->nty;
  a.ntz = c->ntz;
  return a;
}

// One sweep of group g over every octant: q = c->d_q, sigma_t = c->d_sig,
// partials -> c->d_part.  The persistent schedules time themselves through
// c->sweep_ev (events on the kernel dispatch); run_sweep_with records those
// events around the launches of the other schedules.
int
run_sweep_launch(doma_ctx * c, const double * q)
{
  if (c->dims == 1)
  {
    Sweep1DArgs a{};
    a.nx = c->nx;
    a.dirs = c->dirs;
    a.dirtab = c->d_dirtab;
    a.lx = c->d_lx;
    a.q = q;
    a.sig = c->d_sig;
    a.part = c->d_part;
    a.skip = c->si_device ? c->d_si_state : nullptr;
    return launch_sweep1d(a, c->precision, c->stream);
  }
  if (c->schedule == 2)
  {
    Rows2DArgs r{};
    r.nx = c->nx;
    r.ny = c->ny;
    r.N = c->N;
    r.dirs = c->dirs;
    r.dirtab = c->d_dirtab;
    r.lx = c->d_lx;
    r.ly = c->d_ly;
    r.q = q;
    r.sig = c->d_sig;
    r.part = c->d_part;
    r.skip = c->si_device ? c->d_si_state : nullptr;
    r.error = c->d_err;
    r.wait_ticks = (long long)c->wall_clock_khz * 250;  // 250 ms, as the brick waits
    if (c->profiling)
      DOMA_HIP(hipEventRecord(c->prof_events[0], c->stream));
    int rc = launch_sweep2d_rows(r, c->scheme, c->precision, c->exact, c->stream);
    if (c->profiling)
      DOMA_HIP(hipEventRecord(c->prof_events[1], c->stream));
    return rc;
  }
  BrickArgs a = brick_args(c, q);
  a.skip = c->si_device ? c->d_si_state : nullptr;
  if (c->schedule == 1 || c->schedule == 3)
  {
    // Pencil schedule: one task per (pencil, direction-set pair).
    a.ntasks = c->schedule == 3 ? c->npencils * (c->ndirsets / c->ko) : c->ntasks;
    a.pencils = c->d_pencils;
    a.npencils = c->npencils;
    a.flags = c->d_flags;
    a.epoch = ++c->epoch;
    a.ticket = c->d_ticket;
    a.ticket_base = c->ticket_base;
    a.nqueues = c->nqueues;
    a.error = c->d_err;
    // 250 ms per dependency: orders of magnitude above any brick's latency.
    a.wait_ticks = (long long)c->wall_clock_khz * 250;
    // The launch's own dispatch records the timing events (kernel start and end).
    a.ev_start = c->profiling ? c->prof_events[0] : c->sweep_ev[0];
    a.ev_stop = c->profiling ? c->prof_events[1] : c->sweep_ev[1];
    int rc = launch_brick_sweep(a, c->scheme, c->precision, c->exact, c->shape, c->schedule,
                                c->grid_flow, c->stream);
    if (rc)
      return rc;
    // Every work-group takes one ticket past the last task before it exits.
    c->ticket_base += (unsigned long long)a.ntasks + (unsigned long long)c->grid_flow;
    return DOMA_OK;
  }
  const int ndiag = (int)c->diag_off.size() - 1;
  for (int s = 0; s < ndiag; ++s)
  {
    a.diag_begin = c->diag_off[s];
    a.diag_count = c->diag_off[s + 1] - c->diag_off[s];
    if (c->profiling)
      DOMA_HIP(hipEventRecord(c->prof_events[2 * s], c->stream));
    const int grid = a.diag_count * (c->ndirsets / c->ko);
    int rc = launch_brick_sweep(a, c->scheme, c->precision, c->exact, c->shape, 0, grid, c->stream);
    if (rc)
      return rc;
    if (c->profiling)
      DOMA_HIP(hipEventRecord(c->prof_events[2 * s + 1], c->stream));
  }
  return DOMA_OK;
}

int
run_sweep_with(doma_ctx * c, const double * q)
{
  const bool self_timed = c->dims >= 2 && (c->schedule == 1 || c->schedule == 3);
  if (c->sweep_ev[0] && !self_timed)
    DOMA_HIP(hipEventRecord(c->sweep_ev[0], c->stream));
  int rc = run_sweep_launch(c, q);
  if (rc)
    return rc;
  if (c->sweep_ev[1] && !self_timed)
    DOMA_HIP(hipEventRecord(c->sweep_ev[1], c->stream));
  return DOMA_OK;
}

int
run_sweep(doma_ctx * c, int)
{
  return run_sweep_with(c, c->d_q);
}
} // namespace doma

using namespace doma;

namespace
{
int
check_ctx(const doma_ctx * c, bool need_materials)
{
  if (!c)
    return fail(DOMA_ERR_INVALID, "null context");
  if (need_materials && !c->materials)
    return fail(DOMA_ERR_STATE, "materials have not been set (doma_set_materials)");
  return DOMA_OK;
}

// 1D DSA after a sweep: the accelerated flux replaces the swept one in af.
int
enqueue_dsa(doma_ctx * c, IterArgs & af)
{
  if (!(c->dsa && c->dims == 1))
    return DOMA_OK;
  // syntheticAcceleration follows the sweep (TransportSolver1D.cpp:513-514).
  DsaArgs d{};
  d.dirs = c->dirs;
  d.dirtab = c->d_dirtab;
  d.beta_left = c->dsa_beta[0];
  d.beta_right = c->dsa_beta[1];
  d.work = c->d_dsa + c->N;
  d.out = c->d_dsa;
  af.eigen = c->solving_eigen;
  int rc = launch_dsa1d(af, d, c->stream);
  af.dsa = c->d_dsa;
  return rc;
}

// Source iteration with the stopping rule evaluated on the device: the host
// enqueues a chunk of iterations (sweep, DSA, finish, residual) and waits once
// per chunk.  The residual kernel raises a flag after the iteration that meets
// !(tol < r) or it == max_it, and every later kernel of the chunk returns at
// entry, so iterates, iteration count and last residual equal those of one host
// round trip per iteration.  Chunks grow 1, 2, 4, 8 so short iterations enqueue
// few idle launches.
int
si_loop_device(doma_ctx * c, int g, double tol, int max_it, int * its, double * res, int * conv,
               double * sweep_ms, int * sweeps, const double * first_corr)
{
  constexpr int kMaxChunk = 8;
  int rc = DOMA_OK;
  if (c->si_events.empty())
  {
    c->si_events.resize(2);
    for (auto & e : c->si_events)
      DOMA_HIP(hipEventCreate(&e));
  }
  IterArgs a = iter_args(c, g);
  a.si_state = c->d_si_state;
  a.tol = tol;
  a.max_it = max_it;
  if ((rc = launch_fill_i32(c->d_si_state, 4, 0, c->stream)) ||
      (rc = launch_fill(c->d_res + 6, 2, 0.0, c->stream)))
    return rc;
  c->si_device = true;
  struct Off  // sweeps launched after the loop run unconditionally again
  {
    doma_ctx * c;
    ~Off() { c->si_device = false; }
  } off{c};
  int done_its = 0, chunk = 1;
  bool done = false;
  double timed_ms = 0.0;  // duration of the loop's last timed sweep
  while (!done)
  {
    const int n = std::min(chunk, max_it - done_its);
    for (int k = 0; k < n; ++k)
    {
      IterArgs af = a;
      if (done_its + k == 0 && first_corr)
        af.corr = first_corr;  // warm start: the first increment is measured from prev_g
      // Only the loop's first two sweeps are timed (the first of the first two
      // chunks): timing events on a launch cost idle gaps before and after it
      // (about 11 us per sweep at 64^3), so the later sweeps are counted at the
      // second one's duration (the first often runs on cold caches).
      if (k == 0 && done_its <= 1)
      {
        c->sweep_ev[0] = c->si_events[0];
        c->sweep_ev[1] = c->si_events[1];
      }
      rc = run_sweep(c, g);
      c->sweep_ev[0] = c->sweep_ev[1] = nullptr;
      if (rc)
        return rc;
      if ((rc = enqueue_dsa(c, af)) || (rc = launch_finish(af, c->red_blocks, c->stream)) ||
          (rc = launch_residual_final(af, c->red_blocks, c->stream)))
        return rc;
    }
    // res[0..3] as si_loop reads them, res[6..7] = the loop state (the first
    // iteration of every chunk runs, so its residual kernel has written them).
    DOMA_HIP(hipMemcpyAsync(c->h_res, c->d_res, 8 * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    DOMA_HIP(hipStreamSynchronize(c->stream));
    DOMA_HIP(hipGetLastError());
    const int state_done = c->h_res[6] != 0.0;
    const int state_its = static_cast<int>(c->h_res[7]);
    const int ran = state_its - done_its;
    if (done_its <= 1 && ran > 0)
    {
      float ms = 0.f;
      DOMA_HIP(hipEventElapsedTime(&ms, c->si_events[0], c->si_events[1]));
      timed_ms = ms;
    }
    if (sweep_ms && ran > 0)
      *sweep_ms += timed_ms * std::min(ran, n);
    if (sweeps)
      *sweeps += ran;
    done_its = state_its;
    if (c->h_res[3] != 0.0)
      return reset_after_sweep_error(c);
    done = state_done || done_its >= max_it;
    chunk = std::min(2 * chunk, kMaxChunk);
  }
  if (its)
    *its = done_its;
  if (res)
    *res = c->h_res[0];
  if (conv)
    *conv = done_its < max_it ? 1 : 0;
  return DOMA_OK;
}

// Scattering source iteration of group g from the state prepared by the caller
// (TransportSolver3D.cpp:458-497): sweep, residual, source update, until
// it == max_it or the residual is <= tol.
int
si_loop(doma_ctx * c, int g, double tol, int max_it, int * its, double * res, int * conv,
        double * sweep_ms, int * sweeps, doma_stats * st = nullptr,
        const double * first_corr = nullptr)
{
  // With the learned correction every `every` iterations the network's estimate
  // of the remaining error is added to the flux and the next sweep starts from
  // the corrected state; if that sweep's residual is not below the one before
  // the correction, the correction is undone (the sweep still counts).
  const bool learned = c->net.ready && c->net.every > 0 && c->dims == 3;
  const bool err_rule = c->stop_rule == DOMA_STOP_ERROR_ESTIMATE;
  if (!learned && !err_rule && c->si_chunked)
    return si_loop_device(c, g, tol, max_it, its, res, conv, sweep_ms, sweeps, first_corr);
  // Error-estimate rule: rho is the largest ratio of consecutive plain
  // iterations seen so far in this solve for group g -- the ratios approach the
  // spectral radius from below, and after a correction the slow mode's share is
  // reduced for a while, so a recent ratio alone can understate it.
  if ((int)c->rho_est.size() < c->G)
    c->rho_est.assign(c->G, -1.0);
  double r_plain = -1.0;
  auto not_done = [&](double r_) {
    if (!err_rule)
      return tol < r_;
    if (r_ == 0.0)
      return false;
    const double rho = std::min(c->rho_est[g], 0.9999);
    return rho < 0.0 || tol < r_ * rho / (1.0 - rho);
  };
  IterArgs a = iter_args(c, g);
  if (learned)
    a.dcur = c->d_dcur;
  int it = 0;
  double r = 0.0, r_before = 0.0;
  double timed_ms = 0.0;  // duration of the loop's last timed sweep
  bool pending = false;
  // After a rejected correction the next attempt waits twice as long (up to 16x
  // `every`), so a network that does not help costs few wasted sweeps.
  const int every = learned ? c->net.every : 0;
  int gap = every, next_at = every;
  do
  {
    IterArgs af = a;
    if (pending)
      af.corr = c->d_corr;
    else if (it == 0 && first_corr)
      af.corr = first_corr;  // warm start: the first increment is measured from prev_g
    // The loop's first two sweeps are timed and the later ones are counted at
    // the second one's duration (timing events on a launch add idle gaps
    // around it).
    if (it <= 1)
    {
      c->sweep_ev[0] = c->ev0;
      c->sweep_ev[1] = c->ev1;
    }
    int rc = run_sweep(c, g);
    c->sweep_ev[0] = c->sweep_ev[1] = nullptr;
    if (rc)
      return rc;
    if ((rc = enqueue_dsa(c, af)))
      return rc;
    if ((rc = launch_finish(af, c->red_blocks, c->stream)) ||
        (rc = launch_residual_final(af, c->red_blocks, c->stream)))
      return rc;
    DOMA_HIP(hipMemcpyAsync(c->h_res, c->d_res, 4 * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    DOMA_HIP(hipStreamSynchronize(c->stream));
    DOMA_HIP(hipGetLastError());
    if (c->h_res[3] != 0.0)
      return reset_after_sweep_error(c);
    if (it <= 1)
    {
      float ms = 0.f;
      DOMA_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      timed_ms = ms;
    }
    if (sweep_ms)
      *sweep_ms += timed_ms;
    if (sweeps)
      ++*sweeps;
    const double rn = c->h_res[0];
    ++it;
    if (err_rule)
    {
      // A ratio counts only between two plain iterations (this sweep neither
      // follows a correction nor a warm start's first increment).
      const bool plain = !pending && !(it == 1 && first_corr);
      if (plain && r_plain > 0.0)
        c->rho_est[g] = std::max(c->rho_est[g], rn / r_plain);
      r_plain = plain ? rn : -1.0;
    }
    if (!learned)
    {
      r = rn;
      continue;
    }
    if (pending)
    {
      pending = false;
      if (!(rn < r_before))
      {
        IterArgs au = a;
        au.corr = c->d_corr;
        au.dlast = c->d_dlast;
        if ((rc = launch_undo_correction(au, c->d_keep, c->stream)))
          return rc;
        if (st)
          ++st->corrections_rejected;
        r = r_before;
        gap = std::min(2 * gap, 16 * every);
        next_at = it + gap;
        continue;
      }
      if (st)
        ++st->corrections_accepted;
      gap = every;
      next_at = it + gap;
    }
    std::swap(c->d_dlast, c->d_dcur);
    a.dcur = c->d_dcur;
    r = rn;
    if (not_done(r) && it < max_it && it >= next_at)
    {
      const double t0 = std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now().time_since_epoch())
                            .count();
      // d_dlast holds the latest increment and d_dcur the one before: both come
      // from plain iterations, since corrections are at least two sweeps apart.
      if ((rc = net::net_apply(c->net, c->d_dlast, c->d_dcur, c->d_mat, c->d_sigt, c->d_sigs,
                               c->G, g, c->d_corr, c->stream)))
        return rc;
      IterArgs ac = a;
      ac.corr = c->d_corr;
      ac.dlast = c->d_dlast;
      if ((rc = launch_apply_correction(ac, c->d_keep, c->stream)))
        return rc;
      DOMA_HIP(hipStreamSynchronize(c->stream));
      if (st)
        st->correction_ms += std::chrono::duration<double, std::milli>(
                                 std::chrono::steady_clock::now().time_since_epoch())
                                 .count() -
                             t0;
      pending = true;
      r_before = r;
    }
  } while (it < max_it && not_done(r));
  if (its)
    *its = it;
  if (res)
    *res = r;
  if (conv)
    *conv = it < max_it ? 1 : 0;
  return DOMA_OK;
}

double
now_ms()
{
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
} // namespace

extern "C" const char *
doma_last_error(void)
{
  return g_last_error.c_str();
}

extern "C" const char *
doma_version(void)
{
  return "doma_amd 0.1 (gfx950)";
}

namespace doma
{
// Context creation; with a non-null shared stream the context issues its work on
// that stream (batches) and does not own it.
int
create_ctx(const doma_problem_desc * d, hipStream_t shared, doma_ctx ** out)
{
  if (!d || !out)
    return fail(DOMA_ERR_INVALID, "doma_create: null argument");
  *out = nullptr;
  if (d->dims < 1 || d->dims > 3)
    return fail(DOMA_ERR_INVALID, "dims must be 1, 2 or 3");
  const int ny = d->dims > 1 ? d->ny : 1;
  const int nz = d->dims > 2 ? d->nz : 1;
  if (d->nx < 1 || ny < 1 || nz < 1)
    return fail(DOMA_ERR_INVALID, "cell counts must be positive");
  if ((double)d->nx * ny * nz > 2.0e9)
    return fail(DOMA_ERR_INVALID, "mesh too large (cell index must fit in int32)");
  if (d->groups < 1 || d->groups > DOMA_MAX_GROUPS)
    return fail(DOMA_ERR_INVALID, "groups must be in [1, 64]");
  if (d->scheme != DOMA_SCHEME_DD && d->scheme != DOMA_SCHEME_TW_DD)
    return fail(DOMA_ERR_INVALID, "unknown scheme");
  if (d->dims == 1 && d->scheme == DOMA_SCHEME_TW_DD)
    return fail(DOMA_ERR_INVALID,
                "Theta-Weighted Diamond Differences are currently not supported in slab geometries!");
  if (d->precision != DOMA_FP64 && d->precision != DOMA_FP32_SWEEP)
    return fail(DOMA_ERR_INVALID, "unknown precision mode");
  if (d->dirs < 1 || d->dirs > 4096)
    return fail(DOMA_ERR_INVALID, "directions per octant must be in [1, 4096]");
  if (!d->lx || (d->dims > 1 && !d->ly) || (d->dims > 2 && !d->lz) || !d->mu || !d->w ||
      (d->dims > 1 && (!d->eta || !d->xi)))
    return fail(DOMA_ERR_INVALID, "null geometry or quadrature array");
  const int noct = d->dims == 3 ? 8 : (d->dims == 2 ? 4 : 2);
  for (int t = 0; t < noct * d->dirs; ++t)
  {
    if (!(d->w[t] >= 0.0) || !std::isfinite(d->mu[t]) || d->mu[t] == 0.0)
      return fail(DOMA_ERR_INVALID, "invalid direction or weight");
    if (d->dims > 1 && (d->eta[t] == 0.0 || !std::isfinite(d->eta[t]) || !std::isfinite(d->xi[t])))
      return fail(DOMA_ERR_INVALID, "invalid direction");
  }
  auto chk_w = [](const double * l, int n) {
    for (int i = 0; i < n; ++i)
      if (!(l[i] > 0.0) || !std::isfinite(l[i]))
        return false;
    return true;
  };
  if (!chk_w(d->lx, d->nx) || (d->dims > 1 && !chk_w(d->ly, ny)) ||
      (d->dims > 2 && !chk_w(d->lz, nz)))
    return fail(DOMA_ERR_INVALID, "cell widths must be positive and finite");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(DOMA_ERR_DEVICE, "no HIP device available: the doma_amd sweep runs only on the GPU");

  doma_ctx * c = new (std::nothrow) doma_ctx();
  if (!c)
    return fail(DOMA_ERR_INVALID, "out of host memory");
  c->dims = d->dims;
  c->nx = d->nx;
  c->ny = ny;
  c->nz = nz;
  c->N = d->nx * ny * nz;
  c->G = d->groups;
  c->scheme = d->scheme;
  c->precision = d->precision;
  c->dirs = d->dirs;
  c->noct = noct;
  int rc = DOMA_OK;
  auto cleanup = [&](int code) {
    doma_destroy(c);
    return code;
  };
  if (hipGetDevice(&c->device) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "hipGetDevice failed"));
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) == hipSuccess &&
        khz > 0)
      c->wall_clock_khz = khz;
  }
  if (shared)
  {
    c->stream = shared;
    c->own_stream = false;
  }
  else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "stream creation failed"));
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "stream/event creation failed"));

  std::vector<double> lyv(ny, 1.0), lzv(nz, 1.0);
  if (d->dims > 1)
    lyv.assign(d->ly, d->ly + ny);
  if (d->dims > 2)
    lzv.assign(d->lz, d->lz + nz);
  std::vector<double> lxv(d->lx, d->lx + d->nx);
  std::vector<double> rlx(d->nx), rly(ny), rlz(nz);
  for (int i = 0; i < d->nx; ++i)
    rlx[i] = 1.0 / lxv[i];
  for (int j = 0; j < ny; ++j)
    rly[j] = 1.0 / lyv[j];
  for (int k = 0; k < nz; ++k)
    rlz[k] = 1.0 / lzv[k];
  std::vector<double> dirtab((size_t)noct * d->dirs * 4);
  for (int t = 0; t < noct * d->dirs; ++t)
  {
    dirtab[4 * t + 0] = std::fabs(d->mu[t]);
    dirtab[4 * t + 1] = d->dims > 1 ? std::fabs(d->eta[t]) : 0.0;
    dirtab[4 * t + 2] = d->dims > 1 ? std::fabs(d->xi[t]) : 0.0;
    dirtab[4 * t + 3] = d->w[t];
  }
  // 1D DSA boundary coefficients: sum of mu w over each half range, in the
  // reference's order (TransportSolver1D.cpp:763-766, 779-782).
  if (d->dims == 1)
    for (int t = 0; t < noct * d->dirs; ++t)
      c->dsa_beta[t < d->dirs ? 0 : 1] += d->mu[t] * d->w[t];

  const size_t N = (size_t)c->N;
  if ((rc = dev_alloc(&c->d_lx, d->nx)) || (rc = dev_alloc(&c->d_ly, ny)) ||
      (rc = dev_alloc(&c->d_lz, nz)) || (rc = dev_alloc(&c->d_rlx, d->nx)) ||
      (rc = dev_alloc(&c->d_rly, ny)) || (rc = dev_alloc(&c->d_rlz, nz)) ||
      (rc = dev_alloc(&c->d_dirtab, dirtab.size())) || (rc = dev_alloc(&c->d_sig, N)) ||
      (rc = dev_alloc(&c->d_q, N)) || (rc = dev_alloc(&c->d_qmg, N)) ||
      (rc = dev_alloc(&c->d_dphi, N)) || (rc = dev_alloc(&c->d_phi, N * c->G)) ||
      (rc = dev_alloc(&c->d_prev, N * c->G)))
    return cleanup(rc);
  c->red_blocks = reduction_blocks(c->N);
  if ((rc = dev_alloc(&c->d_si_state, 4)))
    return cleanup(rc);
  if ((rc = dev_alloc(&c->d_red, 2 * (size_t)c->red_blocks)) || (rc = dev_alloc(&c->d_res, 8)) ||
      (rc = dev_alloc(&c->d_err, 1)) ||
      (rc = dev_alloc(&c->d_ticket, (size_t)kTicketQueues * kTicketStride)))
    return cleanup(rc);
  if (hipMemset(c->d_err, 0, sizeof(int)) != hipSuccess ||
      hipMemset(c->d_ticket, 0, sizeof(unsigned long long) * kTicketQueues * kTicketStride) !=
          hipSuccess ||
      hipMemset(c->d_res, 0, 8 * sizeof(double)) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "hipMemset failed"));
  if (hipHostMalloc(reinterpret_cast<void **>(&c->h_res), 8 * sizeof(double)) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "hipHostMalloc failed"));
  if (hipMemcpy(c->d_lx, lxv.data(), lxv.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_ly, lyv.data(), lyv.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_lz, lzv.data(), lzv.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_rlx, rlx.data(), rlx.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_rly, rly.data(), rly.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_rlz, rlz.data(), rlz.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_dirtab, dirtab.data(), dirtab.size() * 8, hipMemcpyHostToDevice) !=
          hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "upload of geometry/quadrature failed"));
  if ((rc = launch_fill(c->d_phi, N * c->G, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_prev, N * c->G, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_q, N, 0.0, c->stream)))
    return cleanup(rc);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(tune_key(c));
    if (it != g_tune_cache.end())
    {
      c->tuning = it->second;
      c->autotuned = 1;
    }
  }
  if ((rc = configure_sweep(c)))
    return cleanup(rc);
  if (hipStreamSynchronize(c->stream) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "initialisation failed"));
  *out = c;
  return DOMA_OK;
}
} // namespace doma

extern "C" int
doma_create(const doma_problem_desc * d, doma_ctx ** out)
{
  return create_ctx(d, nullptr, out);
}

extern "C" int
doma_destroy(doma_ctx * c)
{
  if (!c)
    return DOMA_OK;
  if (c->stream)
    (void)hipStreamSynchronize(c->stream);
  free_sweep(c);
  dev_free(c->d_lx);
  dev_free(c->d_ly);
  dev_free(c->d_lz);
  dev_free(c->d_rlx);
  dev_free(c->d_rly);
  dev_free(c->d_rlz);
  dev_free(c->d_dirtab);
  dev_free(c->d_mat);
  dev_free(c->d_sigt);
  dev_free(c->d_sigs);
  dev_free(c->d_src);
  dev_free(c->d_has_src);
  dev_free(c->d_chi);
  dev_free(c->d_prod);
  dev_free(c->d_prompt);
  dev_free(c->d_has_chi);
  dev_free(c->d_has_prod);
  dev_free(c->d_dsa);
  dev_free(c->d_inv_v);
  dev_free(c->d_nd);
  dev_free(c->d_chi_d);
  dev_free(c->d_beta);
  dev_free(c->d_lambda);
  dev_free(c->d_decay);
  dev_free(c->d_step_type);
  dev_free(c->d_step_in);
  dev_free(c->d_step_out);
  dev_free(c->d_step_src);
  dev_free(c->d_last);
  dev_free(c->d_dnp);
  dev_free(c->d_sig);
  dev_free(c->d_q);
  dev_free(c->d_qmg);
  dev_free(c->d_dphi);
  dev_free(c->d_phi);
  dev_free(c->d_prev);
  dev_free(c->d_red);
  dev_free(c->d_res);
  dev_free(c->d_err);
  dev_free(c->d_ticket);
  net::net_free(c->net);
  dev_free(c->d_corr);
  dev_free(c->d_dlast);
  dev_free(c->d_dcur);
  dev_free(c->d_keep);
  if (c->h_res)
    (void)hipHostFree(c->h_res);
  dev_free(c->d_si_state);
  for (auto e : c->si_events)
    (void)hipEventDestroy(e);
  if (c->ev0)
    (void)hipEventDestroy(c->ev0);
  if (c->ev1)
    (void)hipEventDestroy(c->ev1);
  for (auto e : c->prof_events)
    (void)hipEventDestroy(e);
  if (c->stream && c->own_stream)
    (void)hipStreamDestroy(c->stream);
  delete c;
  return DOMA_OK;
}

extern "C" int
doma_set_materials(doma_ctx * c, const int32_t * mat, int32_t M, const double * sigt,
                   const double * sigs, const double * src, const uint8_t * has_src)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  if (!mat || !sigt || !sigs || !src || !has_src || M < 1)
    return fail(DOMA_ERR_INVALID, "doma_set_materials: null array or no materials");
  for (int i = 0; i < c->N; ++i)
    if (mat[i] < 0 || mat[i] >= M)
      return fail(DOMA_ERR_INVALID,
                  "Block " + std::to_string(mat[i]) + " (defined on the mesh) has no material properties!");
  const size_t G = c->G;
  for (size_t e = 0; e < (size_t)M * G; ++e)
    if (!std::isfinite(sigt[e]) || !std::isfinite(src[e]))
      return fail(DOMA_ERR_INVALID, "non-finite cross section or source");
  for (size_t e = 0; e < (size_t)M * G * G; ++e)
    if (!std::isfinite(sigs[e]))
      return fail(DOMA_ERR_INVALID, "non-finite scattering cross section");
  c->fission = false;
  c->kinetics = c->steps_set = c->transient = false;
  if (c->M != M)
  {
    dev_free(c->d_sigt);
    dev_free(c->d_sigs);
    dev_free(c->d_src);
    dev_free(c->d_has_src);
    dev_free(c->d_chi);
    dev_free(c->d_prod);
    dev_free(c->d_prompt);
    dev_free(c->d_has_chi);
    dev_free(c->d_has_prod);
    dev_free(c->d_inv_v);
    dev_free(c->d_nd);
    dev_free(c->d_chi_d);
    dev_free(c->d_beta);
    dev_free(c->d_lambda);
    dev_free(c->d_decay);
    dev_free(c->d_step_type);
    dev_free(c->d_step_in);
    dev_free(c->d_step_out);
    dev_free(c->d_step_src);
    c->M = 0;
    if ((rc = dev_alloc(&c->d_sigt, (size_t)M * G)) || (rc = dev_alloc(&c->d_sigs, (size_t)M * G * G)) ||
        (rc = dev_alloc(&c->d_src, (size_t)M * G)) || (rc = dev_alloc(&c->d_has_src, (size_t)M)))
      return rc;
    c->M = M;
  }
  if (!c->d_mat && (rc = dev_alloc(&c->d_mat, (size_t)c->N)))
    return rc;
  DOMA_HIP(hipMemcpy(c->d_mat, mat, sizeof(int32_t) * c->N, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_sigt, sigt, sizeof(double) * M * G, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_sigs, sigs, sizeof(double) * M * G * G, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_src, src, sizeof(double) * M * G, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_has_src, has_src, (size_t)M, hipMemcpyHostToDevice));
  c->materials = true;
  return DOMA_OK;
}

extern "C" int
doma_set_fission(doma_ctx * c, int32_t M, const double * chi, const double * prod,
                 const double * prompt, const uint8_t * has_chi, const uint8_t * has_prod)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!chi || !prod || !prompt || !has_chi || !has_prod)
    return fail(DOMA_ERR_INVALID, "doma_set_fission: null array");
  if (M != c->M)
    return fail(DOMA_ERR_INVALID, "doma_set_fission: " + std::to_string(M) +
                                      " materials given, doma_set_materials set " +
                                      std::to_string(c->M));
  const size_t G = c->G;
  for (int m = 0; m < M; ++m)
  {
    if (has_chi[m] && !has_prod[m])
      return fail(DOMA_ERR_INVALID, "material " + std::to_string(m) +
                                        " has a prompt fission spectrum but no nu_sigma_f");
    for (size_t g = 0; g < G; ++g)
    {
      const size_t e = (size_t)m * G + g;
      if ((has_chi[m] && (!std::isfinite(chi[e]) || !std::isfinite(prompt[e]))) ||
          (has_prod[m] && !std::isfinite(prod[e])))
        return fail(DOMA_ERR_INVALID, "non-finite fission data");
    }
  }
  if (!c->d_chi)
  {
    if ((rc = dev_alloc(&c->d_chi, (size_t)M * G)) || (rc = dev_alloc(&c->d_prod, (size_t)M * G)) ||
        (rc = dev_alloc(&c->d_prompt, (size_t)M * G)) || (rc = dev_alloc(&c->d_has_chi, (size_t)M)) ||
        (rc = dev_alloc(&c->d_has_prod, (size_t)M)))
      return rc;
  }
  DOMA_HIP(hipStreamSynchronize(c->stream));
  DOMA_HIP(hipMemcpy(c->d_chi, chi, sizeof(double) * M * G, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_prod, prod, sizeof(double) * M * G, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_prompt, prompt, sizeof(double) * M * G, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_has_chi, has_chi, (size_t)M, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_has_prod, has_prod, (size_t)M, hipMemcpyHostToDevice));
  c->fission = true;
  return DOMA_OK;
}

extern "C" int
doma_set_si_mode(doma_ctx * c, int32_t device_loop)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  c->si_chunked = device_loop != 0;
  return DOMA_OK;
}

extern "C" int
doma_set_dsa(doma_ctx * c, int32_t enable)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!enable)
  {
    c->dsa = 0;
    return DOMA_OK;
  }
  if (c->dims != 1)
    return fail(DOMA_ERR_INVALID, "diffusion synthetic acceleration is implemented for 1D only");
  if (c->N < 2)
    return fail(DOMA_ERR_INVALID, "diffusion synthetic acceleration needs at least two cells");
  // Void check of the deck reader (InputParameters.cpp:403-414).
  std::vector<double> sigt((size_t)c->M * c->G);
  DOMA_HIP(hipMemcpy(sigt.data(), c->d_sigt, sizeof(double) * sigt.size(), hipMemcpyDeviceToHost));
  for (double t : sigt)
    if (t < 1e-6)
      return fail(DOMA_ERR_INVALID, "Void region detected! You cannot use synthetic acceleration "
                                    "when cross sections are close to zero.");
  if (!c->d_dsa && (rc = dev_alloc(&c->d_dsa, 5 * (size_t)c->N + 6 * ((size_t)c->N + 1))))
    return rc;
  c->dsa = 1;
  return DOMA_OK;
}

extern "C" int
doma_set_tuning(doma_ctx * c, const doma_tuning * t)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  if (!t)
    return fail(DOMA_ERR_INVALID, "null tuning");
  DOMA_HIP(hipStreamSynchronize(c->stream));
  doma_tuning old = c->tuning;
  c->tuning = *t;
  if ((rc = configure_sweep(c)))
  {
    c->tuning = old;
    int rc2 = configure_sweep(c);
    (void)rc2;
    return rc;
  }
  return DOMA_OK;
}

extern "C" int
doma_get_tuning(const doma_ctx * c, doma_tuning * t)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  if (!t)
    return fail(DOMA_ERR_INVALID, "null tuning");
  t->brick = c->brick;
  t->threads = c->threads;
  t->exact_arith = c->exact;
  t->schedule = c->schedule;
  t->octants = c->ko;
  t->dirs_per_lane = c->ndl;
  t->shape = c->shape + 1;
  return DOMA_OK;
}

extern "C" int
doma_sweep_device(doma_ctx * c, int32_t g, const double * d_q, double * d_dphi)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (g < 0 || g >= c->G || !d_q || !d_dphi)
    return fail(DOMA_ERR_INVALID, "doma_sweep: bad group or null array");
  IterArgs a = iter_args(c, g);
  if ((rc = launch_gather_sig(a, c->stream)) || (rc = run_sweep_with(c, d_q)))
    return rc;
  a.dphi = d_dphi;
  if ((rc = launch_reduce_parts(a, c->stream)))
    return rc;
  DOMA_HIP(hipGetLastError());
  return DOMA_OK;
}

extern "C" int
doma_sweep(doma_ctx * c, int32_t g, const double * q, double * dphi)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!q || !dphi)
    return fail(DOMA_ERR_INVALID, "doma_sweep: null array");
  DOMA_HIP(hipMemcpyAsync(c->d_q, q, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
  if ((rc = doma_sweep_device(c, g, c->d_q, c->d_dphi)))
    return rc;
  DOMA_HIP(hipMemcpyAsync(dphi, c->d_dphi, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
  DOMA_HIP(hipStreamSynchronize(c->stream));
  return check_sweep_error(c);
}

extern "C" int
doma_source_iteration(doma_ctx * c, int32_t g, const double * qmg, double tol, int32_t max_it,
                      double * phi_g, int32_t * iterations, double * residual, int32_t * converged)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (g < 0 || g >= c->G || !qmg)
    return fail(DOMA_ERR_INVALID, "doma_source_iteration: bad group or null source");
  if (!(tol >= 0.0) || max_it < 1)
    return fail(DOMA_ERR_INVALID, "doma_source_iteration: tolerance must be >= 0 and max_it >= 1");
  DOMA_HIP(hipMemcpyAsync(c->d_dphi, qmg, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
  IterArgs a = iter_args(c, g);
  if ((rc = launch_set_q(a, c->d_dphi, c->stream)))
    return rc;
  int its = 0, conv = 0;
  double r = 0.0;
  if ((rc = si_loop(c, g, tol, max_it, &its, &r, &conv, nullptr, nullptr)))
    return rc;
  if (phi_g)
  {
    DOMA_HIP(hipMemcpyAsync(phi_g, c->d_phi + (size_t)g * c->N, sizeof(double) * c->N,
                            hipMemcpyDeviceToHost, c->stream));
    DOMA_HIP(hipStreamSynchronize(c->stream));
  }
  if (iterations)
    *iterations = its;
  if (residual)
    *residual = r;
  if (converged)
    *converged = conv;
  return DOMA_OK;
}

int multigroup_loop(doma_ctx * c, double fs_tol, int fs_max_it, double mg_tol, int mg_max_it,
                    doma_stats & s);

extern "C" int
doma_solve_fixed_source(doma_ctx * c, double fs_tol, int32_t fs_max_it, double mg_tol,
                        int32_t mg_max_it, double * phi, doma_stats * st)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!(fs_tol >= 0.0) || !(mg_tol >= 0.0) || fs_max_it < 1 || mg_max_it < 1)
    return fail(DOMA_ERR_INVALID, "tolerances must be >= 0 and iteration limits >= 1");
  c->transient = false;  // a steady-state solve ends any transient in progress
  c->rho_est.assign(c->G, -1.0);  // contraction estimates belong to one solve
  const double t0 = now_ms();
  const size_t N = (size_t)c->N;
  // initializeSolve (TransportSolver3D.cpp:499-529): phi = 0.
  if ((rc = launch_fill(c->d_phi, N * c->G, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_prev, N * c->G, 0.0, c->stream)))
    return rc;
  doma_stats s{};
  if ((rc = multigroup_loop(c, fs_tol, fs_max_it, mg_tol, mg_max_it, s)))
    return rc;
  if (phi)
  {
    DOMA_HIP(hipMemcpyAsync(phi, c->d_phi, sizeof(double) * N * c->G, hipMemcpyDeviceToHost,
                            c->stream));
  }
  DOMA_HIP(hipStreamSynchronize(c->stream));
  DOMA_HIP(hipGetLastError());
  s.elapsed_ms = now_ms() - t0;
  if (st)
    *st = s;
  return DOMA_OK;
}

// The multigroup Gauss-Seidel loop of solveFixedSource (TransportSolver3D.cpp:
// 272-316) on the current flux state: per pass every group's multigroup source
// and source iteration, then the multigroup residual; stops after mg_max_it
// passes, at a residual <= mg_tol or with one group; success iff fewer than
// mg_max_it passes and every source iteration converged.
int
multigroup_loop(doma_ctx * c, double fs_tol, int fs_max_it, double mg_tol, int mg_max_it,
                doma_stats & s)
{
  int rc = DOMA_OK;
  int mg_it = 0, ok = 1;
  double r = 0.0;
  do
  {
    for (int g = 0; g < c->G; ++g)
    {
      IterArgs a = iter_args(c, g);
      if ((rc = launch_mg_source(a, c->stream)))
        return rc;
      const bool warm = c->warm_start && mg_it > 0 && c->dims > 1;
      if (warm && (rc = launch_warm_start(a, c->stream)))
        return rc;
      int its = 0, conv = 0;
      double res = 0.0;
      if ((rc = si_loop(c, g, fs_tol, fs_max_it, &its, &res, &conv, &s.sweep_ms, &s.sweeps, &s,
                        warm ? c->d_prev + (size_t)g * c->N : nullptr)))
        return rc;
      s.si_iterations[g] = its;
      s.si_residual[g] = res;
      if (!conv)
      {
        ok = 0;
        break;
      }
    }
    if (!ok)
      break;
    IterArgs a = iter_args(c, 0);
    if ((rc = launch_mg_residual(a, c->red_blocks, c->stream)))
      return rc;
    DOMA_HIP(hipMemcpyAsync(c->h_res, c->d_res, 3 * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    DOMA_HIP(hipStreamSynchronize(c->stream));
    r = c->h_res[0];
    ++mg_it;
  } while (mg_it < mg_max_it && mg_tol < r && c->G > 1);
  s.success = ok && mg_it < mg_max_it;
  s.mg_iterations = mg_it;
  s.mg_residual = r;
  return DOMA_OK;
}

extern "C" int
doma_solve_eigenvalue(doma_ctx * c, double fs_tol, int32_t fs_max_it, double mg_tol,
                      int32_t mg_max_it, double k_tol, double * phi, double * k_eff,
                      doma_stats * st)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!c->fission)
    return fail(DOMA_ERR_STATE, "an eigenvalue solve needs fission data (doma_set_fission)");
  if (!(fs_tol >= 0.0) || !(mg_tol >= 0.0) || !(k_tol >= 0.0) || fs_max_it < 1 || mg_max_it < 1)
    return fail(DOMA_ERR_INVALID, "tolerances must be >= 0 and iteration limits >= 1");
  c->transient = false;
  const double t0 = now_ms();
  doma_stats s{};
  const size_t N = (size_t)c->N;
  // initializeSolve, RunMode::Eigen (TransportSolver3D.cpp:499-529): phi = 1,
  // previous iterate and iteration source 0; k = 1, k_prev = 0 (constructor).
  if ((rc = launch_fill(c->d_phi, N * c->G, 1.0, c->stream)) ||
      (rc = launch_fill(c->d_prev, N * c->G, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_q, N, 0.0, c->stream)))
    return rc;
  double k = 1.0, k_prev = 0.0, shape = 0.0, kdiff = 0.0;
  int mg_it = 0, ok = 1;
  struct EigenScope  // the DSA scheme follows the run mode (TransportSolver1D.cpp:603-606)
  {
    doma_ctx * c;
    explicit EigenScope(doma_ctx * x) : c(x) { c->solving_eigen = 1; }
    ~EigenScope() { c->solving_eigen = 0; }
  } scope(c);
  do
  {
    for (int g = 0; g < c->G; ++g)
    {
      IterArgs a = iter_args(c, g);
      a.eigen = 1;
      a.k = k;
      if ((rc = launch_mg_source(a, c->stream)))
        return rc;
      const bool warm = c->warm_start && mg_it > 0 && c->dims > 1;
      if (warm && (rc = launch_warm_start(a, c->stream)))
        return rc;
      int its = 0, conv = 0;
      double res = 0.0;
      if ((rc = si_loop(c, g, fs_tol, fs_max_it, &its, &res, &conv, &s.sweep_ms, &s.sweeps, &s,
                        warm ? c->d_prev + (size_t)g * c->N : nullptr)))
        return rc;
      s.si_iterations[g] = its;
      s.si_residual[g] = res;
      if (!conv)
      {
        ok = 0;
        break;
      }
    }
    if (!ok)
      break;
    // updateEigenvalue, then computeMGFluxResidual (TransportSolver3D.cpp:59-62).
    IterArgs a = iter_args(c, 0);
    a.eigen = 1;
    a.k = k;
    if ((rc = launch_fission_rates(a, c->red_blocks, c->stream)) ||
        (rc = launch_mg_residual(a, c->red_blocks, c->stream)))
      return rc;
    DOMA_HIP(hipMemcpyAsync(c->h_res, c->d_res, 6 * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    DOMA_HIP(hipStreamSynchronize(c->stream));
    k_prev = k;
    k = c->dims == 1 ? k_prev * c->h_res[4] / c->h_res[5] : c->h_res[4] / c->h_res[5];
    shape = c->h_res[0];
    kdiff = std::fabs(k - k_prev);
    ++mg_it;
  } while (mg_it < mg_max_it && (mg_tol < shape || k_tol < kdiff));
  s.success = ok && mg_it < mg_max_it;
  s.mg_iterations = mg_it;
  s.mg_residual = shape;
  s.k_eff = k;
  s.k_residual = kdiff;
  if (phi)
    DOMA_HIP(hipMemcpyAsync(phi, c->d_phi, sizeof(double) * N * c->G, hipMemcpyDeviceToHost,
                            c->stream));
  DOMA_HIP(hipStreamSynchronize(c->stream));
  DOMA_HIP(hipGetLastError());
  s.elapsed_ms = now_ms() - t0;
  if (k_eff)
    *k_eff = k;
  if (st)
    *st = s;
  return DOMA_OK;
}

extern "C" int
doma_set_kinetics(doma_ctx * c, int32_t M, const double * inv_v, const int32_t * num_delayed,
                  const double * chi_d, const double * beta, const double * lambda)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!inv_v || !num_delayed || !chi_d || !beta || !lambda)
    return fail(DOMA_ERR_INVALID, "doma_set_kinetics: null array");
  if (M != c->M)
    return fail(DOMA_ERR_INVALID, "doma_set_kinetics: material count differs from doma_set_materials");
  const size_t G = c->G, K = DOMA_MAX_DELAYED;
  for (int m = 0; m < M; ++m)
  {
    if (num_delayed[m] < 0 || num_delayed[m] > DOMA_MAX_DELAYED)
      return fail(DOMA_ERR_INVALID, "num_delayed must be in [0, " +
                                        std::to_string(DOMA_MAX_DELAYED) + "]");
    for (size_t g = 0; g < G; ++g)
      if (!std::isfinite(inv_v[m * G + g]) || inv_v[m * G + g] < 0.0)
        return fail(DOMA_ERR_INVALID, "inverse velocities must be finite and >= 0");
    for (int d = 0; d < num_delayed[m]; ++d)
    {
      if (!(lambda[m * K + d] > 0.0) || !std::isfinite(lambda[m * K + d]))
        return fail(DOMA_ERR_INVALID, "delayed decay constants must be finite and > 0");
      for (size_t g = 0; g < G; ++g)
        if (!std::isfinite(chi_d[(m * G + g) * K + d]) || !std::isfinite(beta[(m * G + g) * K + d]))
          return fail(DOMA_ERR_INVALID, "non-finite delayed-neutron data");
    }
  }
  if (!c->d_inv_v &&
      ((rc = dev_alloc(&c->d_inv_v, (size_t)M * G)) || (rc = dev_alloc(&c->d_nd, (size_t)M)) ||
       (rc = dev_alloc(&c->d_chi_d, (size_t)M * G * K)) || (rc = dev_alloc(&c->d_beta, (size_t)M * G * K)) ||
       (rc = dev_alloc(&c->d_lambda, (size_t)M * K)) || (rc = dev_alloc(&c->d_decay, (size_t)M * K))))
    return rc;
  DOMA_HIP(hipStreamSynchronize(c->stream));
  DOMA_HIP(hipMemcpy(c->d_inv_v, inv_v, sizeof(double) * M * G, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_nd, num_delayed, sizeof(int32_t) * M, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_chi_d, chi_d, sizeof(double) * M * G * K, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_beta, beta, sizeof(double) * M * G * K, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_lambda, lambda, sizeof(double) * M * K, hipMemcpyHostToDevice));
  c->h_lambda.assign(lambda, lambda + (size_t)M * K);
  c->h_nd.assign(num_delayed, num_delayed + M);
  c->kinetics = true;
  c->transient = false;
  return DOMA_OK;
}

extern "C" int
doma_set_step_sources(doma_ctx * c, int32_t M, const int32_t * type, const double * insert_at,
                      const double * remove_at, const double * src)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!type || !insert_at || !remove_at || !src)
    return fail(DOMA_ERR_INVALID, "doma_set_step_sources: null array");
  if (M != c->M)
    return fail(DOMA_ERR_INVALID,
                "doma_set_step_sources: material count differs from doma_set_materials");
  for (int m = 0; m < M; ++m)
  {
    if (type[m] < DOMA_STEP_NONE || type[m] > DOMA_STEP_REMOVE)
      return fail(DOMA_ERR_INVALID, "unknown step source type");
    if (type[m] == DOMA_STEP_BOTH && remove_at[m] < insert_at[m])
      return fail(DOMA_ERR_INVALID, "The step source for material " + std::to_string(m) +
                                        " is removed before it is inserted!");
  }
  const size_t G = c->G;
  if (!c->d_step_type &&
      ((rc = dev_alloc(&c->d_step_type, (size_t)M)) || (rc = dev_alloc(&c->d_step_in, (size_t)M)) ||
       (rc = dev_alloc(&c->d_step_out, (size_t)M)) || (rc = dev_alloc(&c->d_step_src, (size_t)M * G))))
    return rc;
  DOMA_HIP(hipStreamSynchronize(c->stream));
  DOMA_HIP(hipMemcpy(c->d_step_type, type, sizeof(int32_t) * M, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_step_in, insert_at, sizeof(double) * M, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_step_out, remove_at, sizeof(double) * M, hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(c->d_step_src, src, sizeof(double) * M * G, hipMemcpyHostToDevice));
  c->steps_set = true;
  return DOMA_OK;
}

extern "C" int
doma_transient_begin(doma_ctx * c, double t0, double t1, int32_t steps)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!c->kinetics)
    return fail(DOMA_ERR_STATE, "a transient needs kinetics data (doma_set_kinetics)");
  if (steps < 1 || !std::isfinite(t0) || !std::isfinite(t1) || !(t1 > t0))
    return fail(DOMA_ERR_INVALID, "a transient needs t1 > t0 and steps >= 1");
  for (int m = 0; m < c->M; ++m)
    if (c->h_nd[m] > 0 && !c->fission)
      return fail(DOMA_ERR_STATE, "materials with delayed groups need nu_sigma_f (doma_set_fission)");
  if (c->fission)
  {
    std::vector<uint8_t> has_prod(c->M);
    DOMA_HIP(hipMemcpy(has_prod.data(), c->d_has_prod, (size_t)c->M, hipMemcpyDeviceToHost));
    for (int m = 0; m < c->M; ++m)
      if (c->h_nd[m] > 0 && !has_prod[m])
        return fail(DOMA_ERR_INVALID, "material " + std::to_string(m) +
                                          " has delayed groups but no nu_sigma_f");
  }
  const size_t N = (size_t)c->N, K = DOMA_MAX_DELAYED;
  if (!c->d_last && ((rc = dev_alloc(&c->d_last, N * c->G)) || (rc = dev_alloc(&c->d_dnp, N * K))))
    return rc;
  // dt and exp(-lambda dt) on the host, as the reference computes them
  // (TransportSolver3D.cpp:21, 255).
  c->tr_dt = (t1 - t0) / static_cast<double>(steps);
  std::vector<double> decay((size_t)c->M * K, 0.0);
  for (int m = 0; m < c->M; ++m)
    for (int d = 0; d < c->h_nd[m]; ++d)
      decay[m * K + d] = std::exp(-c->h_lambda[m * K + d] * c->tr_dt);
  DOMA_HIP(hipMemcpy(c->d_decay, decay.data(), sizeof(double) * decay.size(), hipMemcpyHostToDevice));
  // initZeroIC (TransportSolver3D.cpp:140-156).
  if ((rc = launch_fill(c->d_phi, N * c->G, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_prev, N * c->G, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_last, N * c->G, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_dnp, N * K, 0.0, c->stream)) ||
      (rc = launch_fill(c->d_q, N, 0.0, c->stream)))
    return rc;
  c->tr_t = 0.0;
  c->tr_step = 0;
  c->tr_steps = steps;
  c->transient = true;
  DOMA_HIP(hipStreamSynchronize(c->stream));
  return DOMA_OK;
}

extern "C" int
doma_transient_step(doma_ctx * c, double fs_tol, int32_t fs_max_it, double mg_tol,
                    int32_t mg_max_it, double * t_solved, doma_stats * st)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (!c->transient)
    return fail(DOMA_ERR_STATE, "doma_transient_begin has not been called");
  if (c->tr_step >= c->tr_steps)
    return fail(DOMA_ERR_STATE, "all " + std::to_string(c->tr_steps) + " steps have been solved");
  if (!(fs_tol >= 0.0) || !(mg_tol >= 0.0) || fs_max_it < 1 || mg_max_it < 1)
    return fail(DOMA_ERR_INVALID, "tolerances must be >= 0 and iteration limits >= 1");
  const double t0 = now_ms();
  const size_t N = (size_t)c->N;
  // updateStepFluxes between steps (TransportSolver3D.cpp:204-228): the last
  // step's flux becomes phi_last and phi restarts from zero.
  if (c->tr_step > 0)
  {
    DOMA_HIP(hipMemcpyAsync(c->d_last, c->d_phi, sizeof(double) * N * c->G,
                            hipMemcpyDeviceToDevice, c->stream));
    if ((rc = launch_fill(c->d_phi, N * c->G, 0.0, c->stream)))
      return rc;
  }
  doma_stats s{};
  if ((rc = multigroup_loop(c, fs_tol, fs_max_it, mg_tol, mg_max_it, s)))
    return rc;
  if (s.success)
  {
    IterArgs a = iter_args(c, 0);
    if ((rc = launch_step_dnps(a, c->stream)))
      return rc;
  }
  DOMA_HIP(hipStreamSynchronize(c->stream));
  DOMA_HIP(hipGetLastError());
  if (t_solved)
    *t_solved = c->tr_t;
  if (s.success)
  {
    c->tr_t += c->tr_dt;
    ++c->tr_step;
  }
  s.elapsed_ms = now_ms() - t0;
  if (st)
    *st = s;
  return DOMA_OK;
}

extern "C" int
doma_get_dnps(doma_ctx * c, double * dnps, int32_t * nd_max)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  int nd = 0;
  for (int v : c->h_nd)
    nd = std::max(nd, v);
  if (nd_max)
    *nd_max = nd;
  if (!dnps || nd == 0)
    return DOMA_OK;
  if (!c->d_dnp)
    return fail(DOMA_ERR_STATE, "no transient has been started");
  DOMA_HIP(hipMemcpyAsync(dnps, c->d_dnp, sizeof(double) * (size_t)nd * c->N,
                          hipMemcpyDeviceToHost, c->stream));
  DOMA_HIP(hipStreamSynchronize(c->stream));
  return DOMA_OK;
}

extern "C" int
doma_get_flux(doma_ctx * c, double * phi)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  if (!phi)
    return fail(DOMA_ERR_INVALID, "null output");
  DOMA_HIP(hipMemcpyAsync(phi, c->d_phi, sizeof(double) * (size_t)c->N * c->G,
                          hipMemcpyDeviceToHost, c->stream));
  DOMA_HIP(hipStreamSynchronize(c->stream));
  return DOMA_OK;
}

extern "C" int64_t
doma_units_per_sweep(const doma_ctx * c)
{
  if (!c)
    return -1;
  return (int64_t)c->N * c->noct * c->dirs;
}

extern "C" int
doma_time_sweeps(doma_ctx * c, int32_t g, int32_t n, double * ms_per_sweep, int32_t * launches)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (g < 0 || g >= c->G || n < 1)
    return fail(DOMA_ERR_INVALID, "doma_time_sweeps: bad group or count");
  IterArgs a = iter_args(c, g);
  if ((rc = launch_gather_sig(a, c->stream)) || (rc = run_sweep(c, g)))
    return rc;  // warm-up
  DOMA_HIP(hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < n; ++i)
    if ((rc = run_sweep(c, g)))
      return rc;
  DOMA_HIP(hipEventRecord(c->ev1, c->stream));
  DOMA_HIP(hipEventSynchronize(c->ev1));
  DOMA_HIP(hipGetLastError());
  if ((rc = check_sweep_error(c)))
    return rc;
  float ms = 0.f;
  DOMA_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (ms_per_sweep)
    *ms_per_sweep = ms / n;
  if (launches)
    *launches = c->launches_per_sweep;
  return DOMA_OK;
}

extern "C" int
doma_profile_sweeps(doma_ctx * c, int32_t g, int32_t n, double * kernel_ms, double * wall_ms,
                    int32_t * launches)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (g < 0 || g >= c->G || n < 1)
    return fail(DOMA_ERR_INVALID, "doma_profile_sweeps: bad group or count");
  const int L = c->launches_per_sweep;
  const bool per_launch = c->dims != 1;
  if (per_launch)
    while ((int)c->prof_events.size() < 2 * L)
    {
      hipEvent_t e;
      DOMA_HIP(hipEventCreate(&e));
      c->prof_events.push_back(e);
    }
  IterArgs a = iter_args(c, g);
  if ((rc = launch_gather_sig(a, c->stream)) || (rc = run_sweep(c, g)))
    return rc;  // warm-up
  double ksum = 0.0, wsum = 0.0;
  for (int it = 0; it < n; ++it)
  {
    c->profiling = per_launch;
    DOMA_HIP(hipEventRecord(c->ev0, c->stream));
    rc = run_sweep(c, g);
    c->profiling = false;
    if (rc)
      return rc;
    DOMA_HIP(hipEventRecord(c->ev1, c->stream));
    DOMA_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    DOMA_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    wsum += ms;
    if (per_launch)
      for (int s = 0; s < L; ++s)
      {
        DOMA_HIP(hipEventElapsedTime(&ms, c->prof_events[2 * s], c->prof_events[2 * s + 1]));
        ksum += ms;
      }
    else
      ksum += ms;
  }
  DOMA_HIP(hipGetLastError());
  if ((rc = check_sweep_error(c)))
    return rc;
  if (kernel_ms)
    *kernel_ms = ksum / n;
  if (wall_ms)
    *wall_ms = wsum / n;
  if (launches)
    *launches = L;
  return DOMA_OK;
}

// Autotuner objective: n source iterations' worth of work -- each sweep of a
// uniform positive source (scratch[2N..3N), so that every cell carries flux as
// in a converging solve, whatever state the context is in) followed by the
// finish kernel of an iteration (partials summed, residual terms, source
// update) writing its flux and source into scratch[0..2N) -- so a shape pays
// what an iteration pays for its partial planes.  Leaves phi and q untouched.
int
time_iterations(doma_ctx * c, int g, int n, double * ms_per_iteration, double * scratch)
{
  int rc;
  IterArgs a = iter_args(c, g);
  IterArgs f = a;
  f.g = 0;  // the scratch flux holds one group
  f.phi = scratch;
  f.q = scratch + c->N;
  const double * qs = scratch + 2 * (size_t)c->N;
  if ((rc = launch_gather_sig(a, c->stream)) || (rc = run_sweep_with(c, qs)))
    return rc;  // warm-up
  DOMA_HIP(hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < n; ++i)
    if ((rc = run_sweep_with(c, qs)) || (rc = launch_finish(f, c->red_blocks, c->stream)))
      return rc;
  DOMA_HIP(hipEventRecord(c->ev1, c->stream));
  DOMA_HIP(hipEventSynchronize(c->ev1));
  DOMA_HIP(hipGetLastError());
  if ((rc = check_sweep_error(c)))
    return rc;
  float ms = 0.f;
  DOMA_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_per_iteration = ms / n;
  return DOMA_OK;
}

extern "C" int
doma_autotune(doma_ctx * c, int32_t g, doma_tuning * best, double * best_ms)
{
  int rc = check_ctx(c, true);
  if (rc)
    return rc;
  if (g < 0 || g >= c->G)
    return fail(DOMA_ERR_INVALID, "doma_autotune: bad group");
  // Every kernel shape valid for this problem, under each brick schedule.
  std::vector<doma_tuning> cands;
  if (c->dims == 1)
    cands.push_back(doma_tuning{});
  else
  {
    const int minext = c->dims == 3 ? std::min(c->nx, std::min(c->ny, c->nz)) : std::min(c->nx, c->ny);
    for (int s = 0; s < num_shapes(); ++s)
      for (int sch : {0, 1, 3})
      {
        if (sch == 3 && c->dims != 3)
          continue;
        doma_tuning t{};
        t.shape = s + 1;
        t.schedule = sch;
        if (select_shape(c->dims, c->noct, c->dirs, minext, t) == s)
          cands.push_back(t);
      }
    if (c->dims == 2 && rows2d_fits(c->nx, c->ny))
    {
      doma_tuning t{};
      t.schedule = 2;
      cands.push_back(t);
    }
  }
  const doma_tuning keep = c->tuning;
  // Screening: every candidate over 3 iterations.  Final: the best three again,
  // twice each over 8 iterations in round-robin order, the fastest run winning
  // (candidates within a few per cent of each other otherwise swap with noise).
  std::vector<std::pair<double, doma_tuning>> timed;
  double * scratch = nullptr;
  if ((rc = dev_alloc(&scratch, 3 * (size_t)c->N)))
    return rc;
  struct Free
  {
    double * p;
    ~Free() { dev_free(p); }
  } free_scratch{scratch};
  if ((rc = launch_fill(scratch, 2 * (size_t)c->N, 0.0, c->stream)) ||
      (rc = launch_fill(scratch + 2 * (size_t)c->N, c->N, 1.0, c->stream)))
    return rc;
  auto measure = [&](const doma_tuning & t, int n, double * ms) {
    c->tuning = t;
    if (configure_sweep(c) != DOMA_OK)
      return 1;  // shape not available for this direction count
    if ((rc = time_iterations(c, g, n, ms, scratch)))
    {
      c->tuning = keep;
      configure_sweep(c);
      return -1;
    }
    return 0;
  };
  for (const auto & t : cands)
  {
    double ms = 0.0;
    const int r = measure(t, 3, &ms);
    if (r < 0)
      return rc;
    if (r == 0)
      timed.emplace_back(ms, t);
  }
  if (timed.empty())
  {
    c->tuning = keep;
    configure_sweep(c);
    return fail(DOMA_ERR_INVALID, "doma_autotune: no kernel shape fits this problem");
  }
  std::stable_sort(timed.begin(), timed.end(),
                   [](const auto & x, const auto & y) { return x.first < y.first; });
  const size_t finalists = std::min<size_t>(3, timed.size());
  std::vector<double> best_of(finalists, 1e300);
  for (int round = 0; round < 2 && finalists > 1; ++round)
    for (size_t i = 0; i < finalists; ++i)
    {
      double ms = 0.0;
      if (measure(timed[i].second, 8, &ms) < 0)
        return rc;
      best_of[i] = std::min(best_of[i], ms);
    }
  size_t pick = 0;
  for (size_t i = 1; i < finalists; ++i)
    if (best_of[i] < best_of[pick])
      pick = i;
  doma_tuning win = timed[pick].second;
  win.exact_arith = keep.exact_arith;
  const double win_ms = finalists > 1 ? best_of[pick] : timed[0].first;
  c->tuning = win;
  if ((rc = configure_sweep(c)))
    return rc;
  c->autotuned = 1;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[tune_key(c)] = win;
  }
  if (best)
    doma_get_tuning(c, best);
  if (best_ms)
    *best_ms = win_ms;
  return DOMA_OK;
}

extern "C" int64_t
doma_net_param_count(void)
{
  return net::total_params();
}

extern "C" int
doma_set_learned_correction(doma_ctx * c, const float * params, int64_t count, int32_t every)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  if (every < 0 || every == 1)
    return fail(DOMA_ERR_INVALID,
                "doma_set_learned_correction: every must be 0 (off) or >= 2 (the network reads "
                "the last two increments of plain iterations)");
  if (every == 0)
  {
    net::net_free(c->net);
    c->net.every = 0;
    return DOMA_OK;
  }
  if (c->dims != 3)
    return fail(DOMA_ERR_INVALID, "the learned correction is a 3D network");
  if (!params || count != net::total_params())
    return fail(DOMA_ERR_INVALID, "doma_set_learned_correction: expected " +
                                      std::to_string(net::total_params()) + " parameters");
  if ((rc = net::net_setup(c->net, params, c->nz, c->ny, c->nx)))
    return rc;
  const size_t N = (size_t)c->N;
  if (!c->d_corr &&
      ((rc = dev_alloc(&c->d_corr, N)) || (rc = dev_alloc(&c->d_dlast, N)) ||
       (rc = dev_alloc(&c->d_dcur, N)) || (rc = dev_alloc(&c->d_keep, N))))
    return rc;
  c->net.every = every;
  return DOMA_OK;
}

extern "C" int
doma_net_forward(const float * params, int64_t count, int32_t D, int32_t H, int32_t W,
                 const float * features, float * out)
{
  if (!params || !features || !out || count != net::total_params())
    return fail(DOMA_ERR_INVALID, "doma_net_forward: bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(DOMA_ERR_DEVICE, "no HIP device available");
  net::Net n;
  int rc = net::net_setup(n, params, D, H, W);
  if (rc)
  {
    net::net_free(n);
    return rc;
  }
  const size_t V = (size_t)D * H * W;
  float * df = nullptr;
  float * dout = nullptr;
  if (hipMalloc(&df, V * net::CIN * sizeof(float)) != hipSuccess ||
      hipMalloc(&dout, V * sizeof(float)) != hipSuccess)
    rc = fail(DOMA_ERR_DEVICE, "hipMalloc failed");
  if (!rc && hipMemcpy(df, features, V * net::CIN * sizeof(float), hipMemcpyHostToDevice) !=
                 hipSuccess)
    rc = fail(DOMA_ERR_DEVICE, "upload failed");
  if (!rc)
    rc = net::net_forward_features(n, df, dout, nullptr);
  if (!rc && hipMemcpy(out, dout, V * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(DOMA_ERR_DEVICE, "download failed");
  if (df)
    (void)hipFree(df);
  if (dout)
    (void)hipFree(dout);
  net::net_free(n);
  return rc;
}

extern "C" int
doma_set_stopping_rule(doma_ctx * c, int32_t rule)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  if (rule != DOMA_STOP_RESIDUAL && rule != DOMA_STOP_ERROR_ESTIMATE)
    return fail(DOMA_ERR_INVALID, "doma_set_stopping_rule: unknown rule");
  c->stop_rule = rule;
  c->rho_est.assign(c->G, -1.0);
  return DOMA_OK;
}

extern "C" int
doma_set_warm_start(doma_ctx * c, int32_t enable)
{
  int rc = check_ctx(c, false);
  if (rc)
    return rc;
  c->warm_start = enable ? 1 : 0;
  return DOMA_OK;
}

struct doma_net_trainer
{
  net::Trainer * t = nullptr;
};

extern "C" int
doma_net_trainer_create(int32_t crop, int32_t batch, const float * params, int64_t count,
                        doma_net_trainer ** out)
{
  if (!out || !params || count != net::total_params())
    return fail(DOMA_ERR_INVALID, "doma_net_trainer_create: bad arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(DOMA_ERR_DEVICE, "no HIP device available");
  auto * h = new (std::nothrow) doma_net_trainer();
  if (!h)
    return fail(DOMA_ERR_INVALID, "out of host memory");
  const int rc = net::trainer_create(crop, batch, params, &h->t);
  if (rc)
  {
    delete h;
    return rc;
  }
  *out = h;
  return DOMA_OK;
}

extern "C" int
doma_net_trainer_destroy(doma_net_trainer * h)
{
  if (!h)
    return DOMA_OK;
  net::trainer_free(h->t);
  delete h;
  return DOMA_OK;
}

extern "C" int
doma_net_trainer_step(doma_net_trainer * h, const float * features, const float * targets,
                      float lr, float * loss)
{
  if (!h || !features || !targets || !(lr >= 0.f))
    return fail(DOMA_ERR_INVALID, "doma_net_trainer_step: bad arguments");
  return net::trainer_run(h->t, features, targets, lr, true, loss, nullptr);
}

extern "C" int
doma_net_trainer_gradient(doma_net_trainer * h, const float * features, const float * targets,
                          float * loss, float * grad)
{
  if (!h || !features || !targets)
    return fail(DOMA_ERR_INVALID, "doma_net_trainer_gradient: bad arguments");
  return net::trainer_run(h->t, features, targets, 0.f, false, loss, grad);
}

extern "C" int
doma_net_trainer_get_params(doma_net_trainer * h, float * params)
{
  if (!h || !params)
    return fail(DOMA_ERR_INVALID, "doma_net_trainer_get_params: bad arguments");
  return net::trainer_params(h->t, params);
}
