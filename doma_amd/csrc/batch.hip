// batch.hip -- many independent 3D fixed-source problems of uneven size solved
// together: each source-iteration round sweeps every active problem in one
// persistent launch (brick_batch_kernel), then one batched finish/residual pass
// and one small residual copy serve all of them.  Each problem follows the
// reference's solveFixedSource loop exactly (TransportSolver3D.cpp:262-316,
// 458-497), so a problem solved in a batch gives the same iterates as the same
// problem solved alone with the same kernel shape.
//
// Training data: every `sample_every` source iterations of the first multigroup
// pass, the latest two increments (the SI residual history) and the accumulated
// flux of the group are kept on the device; when the group's iteration has converged
// each stored flux becomes the target phi_converged - phi_k (FP32); snapshots
// of iterations that did not converge are dropped.  At most
// `sample_cap` snapshots are kept per (problem, group): when the store is full,
// every other snapshot is dropped and the stride doubles, so the kept snapshots
// stay evenly spread over the iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace doma;

namespace
{
struct SampleSet
{
  int stride = 1;           // snapshots at iterations stride * every * (j + 1)
  int count = 0;
  std::vector<int> its;     // SI iteration of each snapshot
  float * d_res = nullptr;   // [cap][N] latest increment d_k
  float * d_prev = nullptr;  // [cap][N] the increment before it, d_(k-1)
  float * d_stage = nullptr; // [N] d_(k-1), staged one iteration before a snapshot
  double * d_phi = nullptr;  // [cap][N]
  float * d_tgt = nullptr;   // [cap][N]
  bool finalized = false;
};

enum class Phase
{
  Iterating,
  Done
};

struct ProblemState
{
  Phase phase = Phase::Iterating;
  int g = 0;
  int si_it = 0;
  int mg_it = 0;
  doma_stats st{};
};

template <typename T>
int
dalloc(T ** p, size_t n)
{
  *p = nullptr;
  if (n == 0)
    return DOMA_OK;
  DOMA_HIP(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
  return DOMA_OK;
}

template <typename T>
void
dfree(T *& p)
{
  if (p)
    (void)hipFree(p);
  p = nullptr;
}

double
now_ms()
{
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
} // namespace

struct doma_batch
{
  int P = 0;
  int scheme = 0, precision = 0, dirs = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<doma_ctx *> ctx;
  doma_tuning tuning{};
  int shape = -1, pairs = 0, grid = 0;
  bool configured = false;

  BrickArgs * d_probs = nullptr;  // [P]
  uint32_t * d_tasks = nullptr;
  size_t task_cap = 0;
  IterArgs * d_args = nullptr;    // [P] by active slot
  int4 * d_map = nullptr;
  size_t map_cap = 0;
  int * d_nblocks = nullptr;      // [P]
  double * d_res = nullptr;       // [P][4]
  double * h_res = nullptr;       // pinned [P][4]
  unsigned long long * d_ticket = nullptr;  // kTicketQueues counters, kTicketStride apart
  int * d_err = nullptr;
  uint32_t epoch = 0;
  unsigned long long qbase[kTicketQueues] = {};  // each counter's value before the next launch
  int qoff[kTicketQueues + 1] = {};              // the uploaded task list's queue offsets
  int nq = 1;
  int wall_clock_khz = 100000;

  int sample_cap = 0;
  std::vector<std::vector<SampleSet>> samples;  // [P][G]
};

namespace
{
void
free_samples(doma_batch * b)
{
  for (auto & per : b->samples)
    for (auto & s : per)
    {
      dfree(s.d_res);
      dfree(s.d_prev);
      dfree(s.d_stage);
      dfree(s.d_phi);
      dfree(s.d_tgt);
    }
  b->samples.clear();
}

int
check_batch(const doma_batch * b)
{
  if (!b)
    return fail(DOMA_ERR_INVALID, "null batch");
  for (const auto * c : b->ctx)
    if (!c->materials)
      return fail(DOMA_ERR_STATE, "materials have not been set for every problem");
  return DOMA_OK;
}

// Kernel shape for the whole batch: the tuning's shape, or the default for the
// smallest problem's extent (every problem must accept it).
int
configure(doma_batch * b)
{
  b->configured = false;
  int minext = 1 << 30;
  for (const auto * c : b->ctx)
    minext = std::min(minext, std::min(c->nx, std::min(c->ny, c->nz)));
  doma_tuning t = b->tuning;
  t.schedule = 1;
  t.exact_arith = 0;
  const doma_ctx * c0 = b->ctx[0];
  const int shape = select_shape(3, 8, c0->dirs, minext, t);
  if (shape < 0)
    return fail(DOMA_ERR_INVALID, "no 3D sweep kernel shape matches the batch tuning");
  t.shape = shape + 1;
  int rc;
  for (auto * c : b->ctx)
  {
    c->tuning = t;
    if ((rc = configure_sweep(c)))
      return rc;
  }
  b->shape = shape;
  b->pairs = c0->ndirsets / c0->ko;
  if (b->pairs > kBatchMaxPairs)
    return fail(DOMA_ERR_INVALID, "too many direction sets for the batched sweep");
  std::vector<BrickArgs> probs(b->P);
  size_t total = 0;
  for (int p = 0; p < b->P; ++p)
  {
    doma_ctx * c = b->ctx[p];
    const int nbricks = c->ntx * c->nty * c->ntz;
    if (nbricks > kBatchMaxBricks)
      return fail(DOMA_ERR_INVALID, "a problem of the batch has too many bricks");
    probs[p] = brick_args(c, c->d_q);
    probs[p].flags = c->d_flags;
    total += (size_t)nbricks * b->pairs;
  }
  DOMA_HIP(hipMemcpy(b->d_probs, probs.data(), sizeof(BrickArgs) * b->P, hipMemcpyHostToDevice));
  if (total > b->task_cap)
  {
    dfree(b->d_tasks);
    if ((rc = dalloc(&b->d_tasks, total)))
      return rc;
    b->task_cap = total;
  }
  int per_cu = 0;
  if ((rc = batch_blocks_per_cu(b->scheme, b->precision, shape, &per_cu)))
    return rc;
  hipDeviceProp_t prop;
  DOMA_HIP(hipGetDeviceProperties(&prop, b->device));
  b->grid = std::max(1, std::max(per_cu, 1) * prop.multiProcessorCount);
  b->nq = std::max(1, std::min(kTicketQueues, b->grid));
  b->tuning = t;
  b->configured = true;
  return DOMA_OK;
}

// Ticket order of the active problems: by brick diagonal first, then problem,
// brick and direction-set pair, so every task's upwind tasks come earlier; the
// tasks of each ticket queue ((problem * pairs + pair) % nq) are kept in that
// order and the queues are stored one after the other.
int
upload_tasks(doma_batch * b, const std::vector<int> & active, int * ntasks)
{
  std::vector<std::vector<uint32_t>> queues(b->nq);
  int maxdiag = 0;
  for (int p : active)
    maxdiag = std::max(maxdiag, (int)b->ctx[p]->diag_off.size() - 1);
  for (int s = 0; s < maxdiag; ++s)
    for (int p : active)
    {
      const auto & off = b->ctx[p]->diag_off;
      if (s + 1 >= (int)off.size())
        continue;
      for (int br = off[s]; br < off[s + 1]; ++br)
        for (int r = 0; r < b->pairs; ++r)
          queues[(p * b->pairs + r) % b->nq].push_back(((uint32_t)p << 20) |
                                                       ((uint32_t)br << 6) | (uint32_t)r);
    }
  std::vector<uint32_t> tasks;
  for (int q = 0; q < b->nq; ++q)
  {
    b->qoff[q] = (int)tasks.size();
    tasks.insert(tasks.end(), queues[q].begin(), queues[q].end());
  }
  b->qoff[b->nq] = (int)tasks.size();
  *ntasks = (int)tasks.size();
  if (!tasks.empty())
    DOMA_HIP(hipMemcpyAsync(b->d_tasks, tasks.data(), tasks.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, b->stream));
  DOMA_HIP(hipStreamSynchronize(b->stream));
  return DOMA_OK;
}

// Per-slot iterate arguments, block map and block counts of the active problems.
int
upload_iterate(doma_batch * b, const std::vector<int> & active, const std::vector<ProblemState> & ps,
               int * nmap)
{
  std::vector<IterArgs> args;
  std::vector<int4> map;
  std::vector<int> nb;
  for (size_t slot = 0; slot < active.size(); ++slot)
  {
    const int p = active[slot];
    doma_ctx * c = b->ctx[p];
    IterArgs a = iter_args(c, ps[p].g);
    a.res = b->d_res + 4 * slot;
    a.err = b->d_err;
    args.push_back(a);
    nb.push_back(c->red_blocks);
    for (int k = 0; k < c->red_blocks; ++k)
      map.push_back(make_int4((int)slot, k, c->red_blocks, 0));
  }
  if (map.size() > b->map_cap)
  {
    dfree(b->d_map);
    int rc = dalloc(&b->d_map, map.size());
    if (rc)
      return rc;
    b->map_cap = map.size();
  }
  if (!args.empty())
  {
    DOMA_HIP(hipMemcpyAsync(b->d_args, args.data(), args.size() * sizeof(IterArgs),
                            hipMemcpyHostToDevice, b->stream));
    DOMA_HIP(hipMemcpyAsync(b->d_nblocks, nb.data(), nb.size() * sizeof(int),
                            hipMemcpyHostToDevice, b->stream));
    DOMA_HIP(hipMemcpyAsync(b->d_map, map.data(), map.size() * sizeof(int4),
                            hipMemcpyHostToDevice, b->stream));
  }
  DOMA_HIP(hipStreamSynchronize(b->stream));
  *nmap = (int)map.size();
  return DOMA_OK;
}

// One batched sweep of the listed tasks.
int
launch_round(doma_batch * b, int ntasks)
{
  BatchArgs a{};
  a.probs = b->d_probs;
  a.tasks = b->d_tasks;
  a.ntasks = ntasks;
  a.epoch = ++b->epoch;
  a.ticket = b->d_ticket;
  a.nqueues = b->nq;
  for (int q = 0; q < b->nq; ++q)
    a.qbase[q] = b->qbase[q];
  for (int q = 0; q <= b->nq; ++q)
    a.qoff[q] = b->qoff[q];
  a.error = b->d_err;
  a.wait_ticks = (long long)b->wall_clock_khz * 250;
  // A whole number of work-groups per queue, within the resident grid.
  const int grid = std::max(b->nq, std::min(ntasks, b->grid) / b->nq * b->nq);
  int rc = launch_brick_batch(a, b->scheme, b->precision, b->shape, grid, b->stream);
  if (rc)
    return rc;
  // Every work-group takes one ticket past the last task of its queue.
  for (int q = 0; q < b->nq; ++q)
    b->qbase[q] += (unsigned long long)(b->qoff[q + 1] - b->qoff[q]) +
                   (unsigned long long)(grid / b->nq);
  // Keep the contexts' own epochs above every flag value the batch has written.
  for (auto * c : b->ctx)
    c->epoch = b->epoch;
  return DOMA_OK;
}

int
check_error(doma_batch * b)
{
  int e = 0;
  DOMA_HIP(hipMemcpy(&e, b->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (e)
  {
    DOMA_HIP(hipMemset(b->d_err, 0, sizeof(int)));
    return fail(DOMA_ERR_DEVICE, "batched sweep: a dependency wait expired (device stalled)");
  }
  return DOMA_OK;
}

// Start group g of problem p: multigroup source, sigma_t gather (A17).
int
start_group(doma_batch * b, int p, int g)
{
  return launch_mg_source(iter_args(b->ctx[p], g), b->stream);
}

int
capture(doma_batch * b, int p, int g, int it, int every)
{
  SampleSet & s = b->samples[p][g];
  doma_ctx * c = b->ctx[p];
  if (it % (every * s.stride) != 0)
    return DOMA_OK;
  const size_t N = (size_t)c->N;
  if (s.count == b->sample_cap)
  {
    // Drop every other snapshot and double the stride.
    int keep = 0;
    for (int j = 1; j < s.count; j += 2, ++keep)
    {
      DOMA_HIP(hipMemcpyAsync(s.d_res + keep * N, s.d_res + j * N, N * sizeof(float),
                              hipMemcpyDeviceToDevice, b->stream));
      DOMA_HIP(hipMemcpyAsync(s.d_prev + keep * N, s.d_prev + j * N, N * sizeof(float),
                              hipMemcpyDeviceToDevice, b->stream));
      DOMA_HIP(hipMemcpyAsync(s.d_phi + keep * N, s.d_phi + j * N, N * sizeof(double),
                              hipMemcpyDeviceToDevice, b->stream));
      s.its[keep] = s.its[j];
    }
    s.count = keep;
    s.its.resize(keep);
    s.stride *= 2;
    if (it % (every * s.stride) != 0)
      return DOMA_OK;
  }
  if (!s.d_res)
  {
    int rc;
    if ((rc = dalloc(&s.d_res, b->sample_cap * N)) || (rc = dalloc(&s.d_phi, b->sample_cap * N)) ||
        (rc = dalloc(&s.d_prev, b->sample_cap * N)))
      return rc;
  }
  IterArgs a = iter_args(c, g);
  int rc = launch_capture(a, s.d_res + s.count * N, s.d_phi + s.count * N, b->stream);
  if (rc)
    return rc;
  // d_(k-1) was staged after the previous iteration (every iteration before a
  // multiple of `every` stages its increment).
  DOMA_HIP(hipMemcpyAsync(s.d_prev + s.count * N, s.d_stage, N * sizeof(float),
                          hipMemcpyDeviceToDevice, b->stream));
  s.its.push_back(it);
  ++s.count;
  return DOMA_OK;
}

// After iteration `it` of (p, g): keep its increment when the next iteration is a
// multiple of `every` (a snapshot may then need it as d_(k-1)).
int
stage(doma_batch * b, int p, int g, int it, int every)
{
  if ((it + 1) % every != 0)
    return DOMA_OK;
  SampleSet & s = b->samples[p][g];
  doma_ctx * c = b->ctx[p];
  if (!s.d_stage)
  {
    int rc = dalloc(&s.d_stage, (size_t)c->N);
    if (rc)
      return rc;
  }
  double * unused = nullptr;
  return launch_capture(iter_args(c, g), s.d_stage, unused, b->stream);
}

// Targets of (p, g) once its source iteration has ended.  Snapshots of an
// iteration that did not converge have no valid target and are dropped, and so is
// a snapshot taken at the converged iterate itself (its target is zero).
int
finalize_samples(doma_batch * b, int p, int g, bool converged, int final_it)
{
  SampleSet & s = b->samples[p][g];
  doma_ctx * c = b->ctx[p];
  s.finalized = true;
  if (!converged)
    s.count = 0;
  if (s.count > 0 && s.its[s.count - 1] == final_it)
    --s.count;
  s.its.resize(s.count);
  if (s.count == 0)
    return DOMA_OK;
  const size_t N = (size_t)c->N;
  if (!s.d_tgt)
  {
    int rc = dalloc(&s.d_tgt, b->sample_cap * N);
    if (rc)
      return rc;
  }
  return launch_targets(c->d_phi + (size_t)g * N, s.d_phi, s.d_tgt, c->N, s.count, b->stream);
}
} // namespace

extern "C" int
doma_batch_create(int32_t count, const doma_problem_desc * descs, doma_batch ** out)
{
  if (!descs || !out || count < 1)
    return fail(DOMA_ERR_INVALID, "doma_batch_create: need at least one problem");
  *out = nullptr;
  if (count > kBatchMaxProblems)
    return fail(DOMA_ERR_INVALID, "doma_batch_create: at most 4096 problems per batch");
  const doma_problem_desc & d0 = descs[0];
  for (int p = 0; p < count; ++p)
  {
    const doma_problem_desc & d = descs[p];
    if (d.dims != 3)
      return fail(DOMA_ERR_INVALID, "batched problems must be 3D");
    if (d.scheme != d0.scheme || d.precision != d0.precision || d.dirs != d0.dirs)
      return fail(DOMA_ERR_INVALID, "batched problems must share scheme, precision and quadrature");
    if (d.mu && d0.mu)
      for (int t = 0; t < 8 * d.dirs; ++t)
        if (d.mu[t] != d0.mu[t] || d.eta[t] != d0.eta[t] || d.xi[t] != d0.xi[t] ||
            d.w[t] != d0.w[t])
          return fail(DOMA_ERR_INVALID, "batched problems must share the angular quadrature");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(DOMA_ERR_DEVICE, "no HIP device available: the doma_amd sweep runs only on the GPU");
  doma_batch * b = new (std::nothrow) doma_batch();
  if (!b)
    return fail(DOMA_ERR_INVALID, "out of host memory");
  auto cleanup = [&](int code) {
    doma_batch_destroy(b);
    return code;
  };
  b->P = count;
  b->scheme = d0.scheme;
  b->precision = d0.precision;
  b->dirs = d0.dirs;
  if (hipGetDevice(&b->device) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "hipGetDevice failed"));
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess &&
        khz > 0)
      b->wall_clock_khz = khz;
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "stream/event creation failed"));
  b->ctx.assign(count, nullptr);
  int rc;
  for (int p = 0; p < count; ++p)
    if ((rc = create_ctx(&descs[p], b->stream, &b->ctx[p])))
      return cleanup(rc);
  if ((rc = dalloc(&b->d_probs, count)) || (rc = dalloc(&b->d_args, count)) ||
      (rc = dalloc(&b->d_nblocks, count)) || (rc = dalloc(&b->d_res, 4 * (size_t)count)) ||
      (rc = dalloc(&b->d_ticket, (size_t)kTicketQueues * kTicketStride)) ||
      (rc = dalloc(&b->d_err, 1)))
    return cleanup(rc);
  if (hipMemset(b->d_ticket, 0, sizeof(unsigned long long) * kTicketQueues * kTicketStride) !=
          hipSuccess ||
      hipMemset(b->d_err, 0, sizeof(int)) != hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "hipMemset failed"));
  if (hipHostMalloc(reinterpret_cast<void **>(&b->h_res), 4 * sizeof(double) * count) !=
      hipSuccess)
    return cleanup(fail(DOMA_ERR_DEVICE, "hipHostMalloc failed"));
  *out = b;
  return DOMA_OK;
}

extern "C" int
doma_batch_destroy(doma_batch * b)
{
  if (!b)
    return DOMA_OK;
  if (b->stream)
    (void)hipStreamSynchronize(b->stream);
  free_samples(b);
  for (auto * c : b->ctx)
    doma_destroy(c);
  dfree(b->d_probs);
  dfree(b->d_tasks);
  dfree(b->d_args);
  dfree(b->d_map);
  dfree(b->d_nblocks);
  dfree(b->d_res);
  dfree(b->d_ticket);
  dfree(b->d_err);
  if (b->h_res)
    (void)hipHostFree(b->h_res);
  if (b->ev0)
    (void)hipEventDestroy(b->ev0);
  if (b->ev1)
    (void)hipEventDestroy(b->ev1);
  if (b->stream)
    (void)hipStreamDestroy(b->stream);
  delete b;
  return DOMA_OK;
}

extern "C" int
doma_batch_set_materials(doma_batch * b, int32_t p, const int32_t * mat, int32_t M,
                         const double * sigt, const double * sigs, const double * src,
                         const uint8_t * has_src)
{
  if (!b || p < 0 || p >= b->P)
    return fail(DOMA_ERR_INVALID, "doma_batch_set_materials: bad problem index");
  return doma_set_materials(b->ctx[p], mat, M, sigt, sigs, src, has_src);
}

extern "C" int
doma_batch_set_tuning(doma_batch * b, const doma_tuning * t)
{
  if (!b || !t)
    return fail(DOMA_ERR_INVALID, "doma_batch_set_tuning: null argument");
  b->tuning = *t;
  return configure(b);
}

extern "C" int
doma_batch_get_tuning(const doma_batch * b, doma_tuning * t)
{
  if (!b || !t)
    return fail(DOMA_ERR_INVALID, "doma_batch_get_tuning: null argument");
  if (!b->configured)
    return fail(DOMA_ERR_STATE, "the batch has no kernel shape yet");
  return doma_get_tuning(b->ctx[0], t);
}

extern "C" int
doma_batch_time_sweeps(doma_batch * b, int32_t n, double * ms_per_round, int64_t * units)
{
  int rc = check_batch(b);
  if (rc)
    return rc;
  if (n < 1)
    return fail(DOMA_ERR_INVALID, "doma_batch_time_sweeps: n >= 1");
  if (!b->configured && (rc = configure(b)))
    return rc;
  std::vector<int> active(b->P);
  int64_t u = 0;
  for (int p = 0; p < b->P; ++p)
  {
    active[p] = p;
    doma_ctx * c = b->ctx[p];
    // Group 0 with a unit source: the work of a sweep does not depend on the values.
    if ((rc = launch_gather_sig(iter_args(c, 0), b->stream)) ||
        (rc = launch_fill(c->d_q, c->N, 1.0, b->stream)))
      return rc;
    u += (int64_t)c->N * 8 * c->dirs;
  }
  int ntasks = 0;
  if ((rc = upload_tasks(b, active, &ntasks)) || (rc = launch_round(b, ntasks)))
    return rc;
  DOMA_HIP(hipEventRecord(b->ev0, b->stream));
  for (int i = 0; i < n; ++i)
    if ((rc = launch_round(b, ntasks)))
      return rc;
  DOMA_HIP(hipEventRecord(b->ev1, b->stream));
  DOMA_HIP(hipEventSynchronize(b->ev1));
  DOMA_HIP(hipGetLastError());
  if ((rc = check_error(b)))
    return rc;
  float ms = 0.f;
  DOMA_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (ms_per_round)
    *ms_per_round = ms / n;
  if (units)
    *units = u;
  return DOMA_OK;
}

extern "C" int
doma_batch_autotune(doma_batch * b, doma_tuning * best, double * best_ms)
{
  int rc = check_batch(b);
  if (rc)
    return rc;
  const doma_ctx * c0 = b->ctx[0];
  int minext = 1 << 30;
  for (const auto * c : b->ctx)
    minext = std::min(minext, std::min(c->nx, std::min(c->ny, c->nz)));
  const doma_tuning keep = b->tuning;
  doma_tuning win = keep;
  double win_ms = 1e300;
  for (int s = 0; s < num_shapes(); ++s)
  {
    doma_tuning t{};
    t.shape = s + 1;
    t.schedule = 1;
    if (select_shape(3, 8, c0->dirs, minext, t) != s)
      continue;
    b->tuning = t;
    if (configure(b) != DOMA_OK)
      continue;
    double ms = 0.0;
    if ((rc = doma_batch_time_sweeps(b, 2, &ms, nullptr)))
    {
      b->tuning = keep;
      configure(b);
      return rc;
    }
    if (ms < win_ms)
    {
      win_ms = ms;
      win = t;
    }
  }
  b->tuning = win;
  if ((rc = configure(b)))
    return rc;
  if (best)
    doma_batch_get_tuning(b, best);
  if (best_ms)
    *best_ms = win_ms;
  return DOMA_OK;
}

extern "C" int
doma_batch_solve(doma_batch * b, double fs_tol, int32_t fs_max_it, double mg_tol,
                 int32_t mg_max_it, int32_t sample_every, int32_t sample_cap, doma_stats * stats,
                 doma_batch_info * info)
{
  int rc = check_batch(b);
  if (rc)
    return rc;
  if (!(fs_tol >= 0.0) || !(mg_tol >= 0.0) || fs_max_it < 1 || mg_max_it < 1)
    return fail(DOMA_ERR_INVALID, "tolerances must be >= 0 and iteration limits >= 1");
  if (sample_every < 0 || sample_every == 1 || sample_cap < 0 ||
      (sample_every > 0 && sample_cap < 2) ||
      sample_cap > 4096)
    return fail(DOMA_ERR_INVALID,
                "sampling: every = 0 (off) or >= 2, and 2 <= cap <= 4096 when sampling");
  if (!b->configured && (rc = configure(b)))
    return rc;
  const double t0 = now_ms();
  free_samples(b);
  b->sample_cap = sample_every > 0 ? sample_cap : 0;
  b->samples.assign(b->P, {});
  for (int p = 0; p < b->P; ++p)
    b->samples[p].assign(b->ctx[p]->G, SampleSet{});

  std::vector<ProblemState> ps(b->P);
  for (int p = 0; p < b->P; ++p)
  {
    doma_ctx * c = b->ctx[p];
    const size_t N = (size_t)c->N;
    if ((rc = launch_fill(c->d_phi, N * c->G, 0.0, b->stream)) ||
        (rc = launch_fill(c->d_prev, N * c->G, 0.0, b->stream)) || (rc = start_group(b, p, 0)))
      return rc;
  }
  doma_batch_info bi{};
  std::vector<int> active(b->P);
  for (int p = 0; p < b->P; ++p)
    active[p] = p;
  bool changed = true;
  int ntasks = 0, nmap = 0;
  double sweep_ms = 0.0;
  while (!active.empty())
  {
    if (changed)
    {
      if ((rc = upload_tasks(b, active, &ntasks)) || (rc = upload_iterate(b, active, ps, &nmap)))
        return rc;
      changed = false;
    }
    DOMA_HIP(hipEventRecord(b->ev0, b->stream));
    if ((rc = launch_round(b, ntasks)))
      return rc;
    DOMA_HIP(hipEventRecord(b->ev1, b->stream));
    if ((rc = launch_batch_finish(b->d_args, b->d_map, nmap, b->stream)) ||
        (rc = launch_batch_residual(b->d_args, b->d_nblocks, (int)active.size(), b->stream)))
      return rc;
    DOMA_HIP(hipMemcpyAsync(b->h_res, b->d_res, 4 * sizeof(double) * active.size(),
                            hipMemcpyDeviceToHost, b->stream));
    DOMA_HIP(hipStreamSynchronize(b->stream));
    DOMA_HIP(hipGetLastError());
    ++bi.rounds;
    for (int p : active)
      bi.units += (int64_t)b->ctx[p]->N * 8 * b->ctx[p]->dirs;
    float ms = 0.f;
    DOMA_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    sweep_ms += ms;
    for (size_t slot = 0; slot < active.size(); ++slot)
      if (b->h_res[4 * slot + 3] != 0.0)
      {
        const int e = check_error(b);
        return e ? e : fail(DOMA_ERR_DEVICE, "batched sweep failed");
      }

    std::vector<int> next;
    for (size_t slot = 0; slot < active.size(); ++slot)
    {
      const int p = active[slot];
      ProblemState & s = ps[p];
      doma_ctx * c = b->ctx[p];
      const double r = b->h_res[4 * slot];
      ++s.si_it;
      ++s.st.sweeps;
      if (sample_every > 0 && s.mg_it == 0 &&
          ((rc = capture(b, p, s.g, s.si_it, sample_every)) ||
           (rc = stage(b, p, s.g, s.si_it, sample_every))))
        return rc;
      if (s.si_it < fs_max_it && fs_tol < r)
      {
        next.push_back(p);
        continue;
      }
      // Source iteration of group g has ended (TransportSolver3D.cpp:458-497).
      changed = true;
      s.st.si_iterations[s.g] = s.si_it;
      s.st.si_residual[s.g] = r;
      if (sample_every > 0 && s.mg_it == 0 &&
          (rc = finalize_samples(b, p, s.g, s.si_it < fs_max_it, s.si_it)))
        return rc;
      if (s.si_it >= fs_max_it)
      {
        s.st.success = 0;
        s.phase = Phase::Done;
        continue;
      }
      s.si_it = 0;
      if (++s.g < c->G)
      {
        if ((rc = start_group(b, p, s.g)))
          return rc;
        next.push_back(p);
        continue;
      }
      // End of a multigroup pass (TransportSolver3D.cpp:292-316).
      IterArgs a = iter_args(c, 0);
      if ((rc = launch_mg_residual(a, c->red_blocks, b->stream)))
        return rc;
      DOMA_HIP(hipMemcpyAsync(c->h_res, c->d_res, 3 * sizeof(double), hipMemcpyDeviceToHost,
                              b->stream));
      DOMA_HIP(hipStreamSynchronize(b->stream));
      const double rmg = c->h_res[0];
      ++s.mg_it;
      s.st.mg_iterations = s.mg_it;
      s.st.mg_residual = rmg;
      if (s.mg_it < mg_max_it && mg_tol < rmg && c->G > 1)
      {
        s.g = 0;
        if ((rc = start_group(b, p, 0)))
          return rc;
        next.push_back(p);
        continue;
      }
      s.st.success = s.mg_it < mg_max_it ? 1 : 0;
      s.phase = Phase::Done;
    }
    active.swap(next);
  }
  DOMA_HIP(hipStreamSynchronize(b->stream));
  DOMA_HIP(hipGetLastError());
  bi.elapsed_ms = now_ms() - t0;
  bi.sweep_ms = sweep_ms;
  for (int p = 0; p < b->P; ++p)
  {
    ps[p].st.elapsed_ms = bi.elapsed_ms;
    if (stats)
      stats[p] = ps[p].st;
    for (const auto & s : b->samples[p])
      bi.samples += s.count;
  }
  if (info)
    *info = bi;
  return DOMA_OK;
}

extern "C" int
doma_batch_get_flux(doma_batch * b, int32_t p, double * phi)
{
  if (!b || p < 0 || p >= b->P || !phi)
    return fail(DOMA_ERR_INVALID, "doma_batch_get_flux: bad argument");
  return doma_get_flux(b->ctx[p], phi);
}

extern "C" int
doma_batch_sample_count(const doma_batch * b, int32_t p, int32_t g, int32_t * count,
                        int32_t * iterations)
{
  if (!b || p < 0 || p >= b->P || !count)
    return fail(DOMA_ERR_INVALID, "doma_batch_sample_count: bad argument");
  if (b->samples.empty() || g < 0 || g >= b->ctx[p]->G)
  {
    *count = 0;
    return g < 0 || g >= b->ctx[p]->G ? fail(DOMA_ERR_INVALID, "bad group") : DOMA_OK;
  }
  const SampleSet & s = b->samples[p][g];
  *count = s.finalized ? s.count : 0;
  if (iterations && s.finalized)
    for (int k = 0; k < s.count; ++k)
      iterations[k] = s.its[k];
  return DOMA_OK;
}

extern "C" int
doma_batch_get_sample(doma_batch * b, int32_t p, int32_t g, int32_t k, float * residual,
                      float * prev_residual, float * target)
{
  if (!b || p < 0 || p >= b->P || b->samples.empty() || g < 0 || g >= b->ctx[p]->G)
    return fail(DOMA_ERR_INVALID, "doma_batch_get_sample: bad argument");
  const SampleSet & s = b->samples[p][g];
  if (!s.finalized || k < 0 || k >= s.count)
    return fail(DOMA_ERR_INVALID, "doma_batch_get_sample: no such sample");
  const size_t N = (size_t)b->ctx[p]->N;
  if (residual)
    DOMA_HIP(hipMemcpy(residual, s.d_res + k * N, N * sizeof(float), hipMemcpyDeviceToHost));
  if (prev_residual)
    DOMA_HIP(hipMemcpy(prev_residual, s.d_prev + k * N, N * sizeof(float),
                       hipMemcpyDeviceToHost));
  if (target)
    DOMA_HIP(hipMemcpy(target, s.d_tgt + k * N, N * sizeof(float), hipMemcpyDeviceToHost));
  return DOMA_OK;
}
