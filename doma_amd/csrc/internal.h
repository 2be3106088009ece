// internal.h -- host-side context and helpers shared by the C-ABI translation
// units.  Not installed; no symbol here is exported.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "doma_c.h"
#include "unet.h"

namespace doma
{
int fail(int code, const std::string & msg);
int hip_fail(hipError_t e, const char * what);

#define DOMA_HIP(call)                                                                           \
  do                                                                                             \
  {                                                                                              \
    hipError_t _e = (call);                                                                      \
    if (_e != hipSuccess)                                                                        \
      return ::doma::hip_fail(_e, #call);                                                        \
  } while (0)

// Ticket queues of the dataflow sweep: at most kTicketQueues counters, each on
// its own 256-byte line (kTicketStride counters apart).
constexpr int kTicketQueues = 16;
constexpr int kTicketStride = 32;

// Arguments of the brick-wavefront sweep kernel (2D and 3D).
struct BrickArgs
{
  int nx, ny, nz;          // cells per axis
  int N;                   // nx*ny*nz
  int dirs;                // directions per octant
  int chunk;               // directions per direction set (<= DL)
  int nchunks;             // direction sets per octant
  int ndirsets;            // octants * nchunks
  const double * dirtab;   // [noct*dirs][4] = |mu|, |eta|, |xi|, w
  const double * lx;       // [nx]
  const double * ly;       // [ny]
  const double * lz;       // [nz]
  const double * rlx;      // [nx] 1/lx (fast DD)
  const double * rly;
  const double * rlz;
  const double * q;        // [N] iteration source
  const double * sig;      // [N] sigma_t of the group being swept
  double * part;           // [ndirsets][N] per-direction-set partial scalar flux
  void * fx;               // global face buffers (scalar type of the sweep)
  void * fy;
  void * fz;
  int nxp, nyp, nzp;       // extents padded to whole bricks
  int ntx, nty, ntz;       // bricks per axis
  const int16_t * bricks;  // [nbricks][3] oriented brick coordinates, sorted by diagonal
  int diag_begin;          // first brick of this launch's diagonal
  int diag_count;          // bricks on the diagonal
  const uint32_t * plane_cells;  // brick-local cells ordered by plane (li | lj<<8 | lk<<16)
  const int32_t * plane_off;     // [nplanes+1]

  // Dataflow schedule (one persistent launch per sweep).
  int ntasks;                       // nbricks * ndirsets, ticket order = diagonal order
  uint32_t * flags;                 // [ndirsets][nbricks] epoch of the last finished brick
  uint32_t epoch;                   // this sweep's epoch (> every flag value before it)
  // Tickets come from nqueues monotone counters, kTicketStride apart: work-group
  // b takes the tasks of queue b % nqueues, whose direction-set pairs are those
  // r with r % nqueues == queue.  nqueues divides the pairs per brick and the
  // grid, so every queue holds ntasks / nqueues tasks and grid / nqueues groups.
  unsigned long long * ticket;      // [nqueues * kTicketStride], never reset mid-run
  unsigned long long ticket_base;   // sum of every queue's counter at the sweep's start
  int nqueues;
  // Pencil schedule: a task is a column of bricks along z, swept bk = 0, 1, ...
  // by one work-group that keeps the z faces in LDS.
  const int16_t * pencils;          // [npencils][2] oriented (bi, bj), sorted by bi+bj
  int npencils;
  int * error;                      // set when a bounded wait expires
  long long wait_ticks;             // per-dependency deadline in device wall-clock ticks
  // Host side only: events recorded by the persistent launch's own dispatch
  // (its start and end), so that timing a sweep adds no marker to the stream.
  hipEvent_t ev_start;
  hipEvent_t ev_stop;

  // Device-side source iteration: when *skip != 0 (the iteration has converged)
  // the launch does nothing (null: always run).  A skipped dataflow launch still
  // advances the ticket counters by ntasks + grid in total, as a full one does.
  const int * skip;
};

// Arguments of the batched persistent sweep: many 3D problems swept with one
// kernel shape in one launch.  A task code is (problem << 20) | (brick << 6) |
// pair, where brick indexes the problem's diagonal-sorted brick list and pair =
// octant group * nchunks + chunk.  Tasks are split over nqueues ticket queues by
// (problem * pairs + pair) % nqueues, so a task and its upwind tasks share a
// queue; queue q's tasks are tasks[qoff[q] .. qoff[q+1]) in ticket order, served
// by the work-groups b with b % nqueues == q.
struct BatchArgs
{
  const BrickArgs * probs;         // [P] per-problem sweep arguments (device)
  const uint32_t * tasks;          // [ntasks], grouped by queue
  int ntasks;
  uint32_t epoch;                  // > every flag value of every problem before it
  unsigned long long * ticket;     // [kTicketQueues * kTicketStride] monotone counters
  unsigned long long qbase[kTicketQueues];  // each queue's counter at the launch's start
  int qoff[kTicketQueues + 1];
  int nqueues;
  int * error;                     // set when a bounded wait expires
  long long wait_ticks;            // per-dependency deadline in wall-clock ticks
};
constexpr int kBatchMaxProblems = 1 << 12;
constexpr int kBatchMaxBricks = 1 << 14;
constexpr int kBatchMaxPairs = 1 << 6;

struct Sweep1DArgs
{
  int nx;
  int dirs;
  const double * dirtab;  // [2*dirs][4]
  const double * lx;
  const double * q;
  const double * sig;
  double * part;          // [2*dirs][nx]
  const int * skip;       // as BrickArgs::skip
};

// 2D sweep with one lane per mesh row (sweep_rows.hip), schedule 2.
constexpr int kRowsMaxThreads = 512;  // rows per work-group (one lane each)
struct Rows2DArgs
{
  int nx, ny, N;
  int dirs;                // directions per quadrant
  const double * dirtab;   // [4*dirs][4]: |mu|, |eta|, |xi|, w
  const double * lx;       // [nx]
  const double * ly;       // [ny]
  const double * q;        // [N]
  const double * sig;      // [N]
  double * part;           // [4*dirs][N] per-direction w psi
  const int * skip;        // as BrickArgs::skip
  int * error;             // set if a wave-to-wave wait exceeds its bound
  long long wait_ticks;    // that bound, in device wall-clock ticks
};
size_t rows2d_lds(int nx, int ny, size_t rs);
bool rows2d_fits(int nx, int ny);
int launch_sweep2d_rows(const Rows2DArgs & a, int scheme, int precision, int exact, hipStream_t s);

// Kernel shapes of the brick sweep (table in sweep_impl.h).
struct ShapeInfo
{
  int dims, brick, dl, ndl, ko, threads;
};
bool shape_info(int shape, ShapeInfo * out);
int num_shapes();
size_t shape_lds_bytes(int shape, int precision);

// Launch wrappers implemented in sweep.hip.  schedule 0: one launch per brick
// diagonal (grid = work-groups of this diagonal); 1: persistent dataflow launch.
int launch_brick_sweep(const BrickArgs & a, int scheme, int precision, int exact, int shape,
                       int schedule, int grid, hipStream_t s);
int sweep_blocks_per_cu(int scheme, int precision, int exact, int shape, int schedule,
                        int * per_cu);
int launch_sweep1d(const Sweep1DArgs & a, int precision, hipStream_t s);
// Batched persistent sweep (3D shapes only, fast arithmetic).
int launch_brick_batch(const BatchArgs & b, int scheme, int precision, int shape, int grid,
                       hipStream_t s);
int batch_blocks_per_cu(int scheme, int precision, int shape, int * per_cu);

// Launch wrappers implemented in iterate.hip.
struct IterArgs
{
  int dims;
  int N;
  int nx, ny;
  int G;
  int g;
  int ndirsets;
  const int32_t * mat;
  const double * sigt;     // [M][G]
  const double * sigs;     // [M][G][G]
  const double * src;      // [M][G]
  const uint8_t * has_src; // [M]
  const double * lx;
  const double * ly;
  const double * lz;
  double * sig;            // [N] gathered sigma_t of group g
  double * q;              // [N]
  double * qmg;            // [N] (1D multigroup source)
  double * phi;            // [G][N]
  double * prev;           // [G][N]
  double * part;           // [ndirsets][N]
  double * dphi;           // [N] (used by the plain sweep entry point)
  double * red;            // reduction partials
  double * res;            // [8] residual, diff, total, sweep error flag; [4..5] fission
                           // rates (power iteration); [6..7] device SI loop state copy
  const int * err;         // sweep error flag (bounded wait expired)
  const double * corr;     // [N] correction applied before this sweep (or null)
  double * dcur;           // [N] receives this iteration's increment (or null)
  const double * dlast;    // [N] last accepted increment (correction kernels)
  const double * chi;      // [M][G] prompt fission spectrum (null: no fission data)
  const double * prod;     // [M][G] nu sigma_f
  const double * prompt;   // [M][G] prompt fraction 1 - sum_d beta
  const uint8_t * has_chi; // [M]
  const uint8_t * has_prod;// [M]
  int eigen;               // 1: power-iteration source (fission / k, no external source)
  double k;                // current eigenvalue (eigen != 0)
  const double * dsa;      // [N] 1D: accelerated flux replacing the swept one (or null)
  // Transient step (null inv_v: steady state).
  const double * inv_v;    // [M][G]
  const int32_t * nd;      // [M] delayed groups per material
  const double * chi_d;    // [M][G][DOMA_MAX_DELAYED]
  const double * beta;     // [M][G][DOMA_MAX_DELAYED]
  const double * lambda;   // [M][DOMA_MAX_DELAYED]
  const double * decay;    // [M][DOMA_MAX_DELAYED] exp(-lambda dt)
  const int32_t * step_type;  // [M] DOMA_STEP_* (null: no step sources)
  const double * step_in;  // [M]
  const double * step_out; // [M]
  const double * step_src; // [M][G]
  const double * last;     // [G][N] flux of the previous step
  double * dnp;            // [DOMA_MAX_DELAYED][N] precursor concentrations
  double t;                // time of the step
  double dt;
  // Device-side source iteration (null si_state: one host round trip per
  // iteration).  si_state[0] = converged-or-done flag, si_state[1] = iterations
  // done; the residual kernel sets the flag once !(tol < r), the count reaches
  // max_it or a sweep reported an expired wait, and every later kernel of the
  // iteration returns at entry.
  int * si_state;
  double tol;
  int max_it;
};

// 1D diffusion synthetic acceleration of one sweep (see iterate.hip).
struct DsaArgs
{
  int dirs;                // directions per half range
  const double * dirtab;   // [2*dirs][4]: |mu|, -, -, w
  double beta_left;        // sum_{mu > 0} mu w
  double beta_right;       // sum_{mu < 0} mu w
  double * work;           // [4 N + 6 (N + 1)] moments, system, solution
  double * out;            // [N] accelerated scalar flux
};

int launch_mg_source(const IterArgs & a, hipStream_t s);
int launch_gather_sig(const IterArgs & a, hipStream_t s);
int launch_finish(const IterArgs & a, int nblocks, hipStream_t s);
int launch_residual_final(const IterArgs & a, int nblocks, hipStream_t s);
int launch_reduce_parts(const IterArgs & a, hipStream_t s);
int launch_mg_residual(const IterArgs & a, int nblocks, hipStream_t s);
// Fission production sums of updateEigenvalue: res[4] = sum V nuSf phi,
// res[5] = sum V nuSf prev (/ k in 2D/3D).  Uses the reduction partials.
int launch_fission_rates(const IterArgs & a, int nblocks, hipStream_t s);
// 1D DSA after a sweep: moments from the direction partials, the tridiagonal
// diffusion system, its solve, the cell-average flux into d.out.
int launch_dsa1d(const IterArgs & a, const DsaArgs & d, hipStream_t s);
// stepDNPs (TransportSolver3D.cpp:230-260): precursor update after a step.
int launch_step_dnps(const IterArgs & a, hipStream_t s);
int launch_fill(double * p, size_t n, double v, hipStream_t s);
int launch_fill_i32(int * p, size_t n, int v, hipStream_t s);
// Batched finish/residual over problem slots (see iterate.hip), training-data
// snapshots and targets.
int launch_batch_finish(const IterArgs * d_args, const int4 * d_map, int nblocks, hipStream_t s);
int launch_batch_residual(const IterArgs * d_args, const int * d_nblocks, int nprob,
                          hipStream_t s);
int launch_capture(const IterArgs & a, float * res_out, double * phi_out, hipStream_t s);
int launch_targets(const double * phi_g, const double * snaps, float * targets, int N, int count,
                   hipStream_t s);
// Learned correction: keep = phi_g, phi_g += corr, q = S (dlast + corr); undo:
// phi_g = keep, q = S dlast.
int launch_apply_correction(const IterArgs & a, double * keep, hipStream_t s);
int launch_undo_correction(const IterArgs & a, const double * keep, hipStream_t s);
// Warm start: phi_g = prev_g, q += S prev_g (2D/3D).
int launch_warm_start(const IterArgs & a, hipStream_t s);
int launch_set_q(const IterArgs & a, const double * qmg_dev, hipStream_t s);
int reduction_blocks(int N);
} // namespace doma

struct doma_ctx
{
  int dims = 0, nx = 0, ny = 0, nz = 0, G = 0, scheme = 0, precision = 0, dirs = 0;
  int noct = 0;
  int N = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // geometry / quadrature (device)
  double * d_lx = nullptr;
  double * d_ly = nullptr;
  double * d_lz = nullptr;
  double * d_rlx = nullptr;
  double * d_rly = nullptr;
  double * d_rlz = nullptr;
  double * d_dirtab = nullptr;

  // materials (device)
  int M = 0;
  bool materials = false;
  int32_t * d_mat = nullptr;
  double * d_sigt = nullptr;
  double * d_sigs = nullptr;
  double * d_src = nullptr;
  uint8_t * d_has_src = nullptr;
  bool fission = false;  // doma_set_fission data present for the current materials
  int dsa = 0;            // 1D diffusion synthetic acceleration (doma_set_dsa)
  // Transient state (doma_set_kinetics, doma_set_step_sources, doma_transient_*).
  bool kinetics = false, steps_set = false, transient = false;
  double * d_inv_v = nullptr;
  int32_t * d_nd = nullptr;
  double * d_chi_d = nullptr;
  double * d_beta = nullptr;
  double * d_lambda = nullptr;
  double * d_decay = nullptr;
  int32_t * d_step_type = nullptr;
  double * d_step_in = nullptr;
  double * d_step_out = nullptr;
  double * d_step_src = nullptr;
  double * d_last = nullptr;  // [G][N]
  double * d_dnp = nullptr;   // [DOMA_MAX_DELAYED][N]
  std::vector<double> h_lambda;  // [M][DOMA_MAX_DELAYED]
  std::vector<int32_t> h_nd;
  double tr_t = 0.0, tr_dt = 0.0;
  int tr_step = 0, tr_steps = 0;
  int solving_eigen = 0;  // inside doma_solve_eigenvalue (selects the DSA scheme)
  double dsa_beta[2] = {0.0, 0.0};
  double * d_dsa = nullptr;  // DSA work [4 N + 6 (N + 1)] + output [N]
  double * d_chi = nullptr;
  double * d_prod = nullptr;
  double * d_prompt = nullptr;
  uint8_t * d_has_chi = nullptr;
  uint8_t * d_has_prod = nullptr;

  // iteration state (device)
  double * d_sig = nullptr;
  double * d_q = nullptr;
  double * d_qmg = nullptr;
  double * d_dphi = nullptr;
  double * d_phi = nullptr;
  double * d_prev = nullptr;
  double * d_part = nullptr;
  double * d_red = nullptr;
  double * d_res = nullptr;
  double * h_res = nullptr;  // pinned
  int red_blocks = 0;

  // sweep decomposition
  doma_tuning tuning{};
  int shape = -1;  // kernel shape id (sweep_impl.h)
  int brick = 0, dl = 0, threads = 0, exact = 0, ndl = 1, ko = 1, dw = 0;
  int chunk = 0, nchunks = 0, ndirsets = 0;
  int nxp = 0, nyp = 0, nzp = 0;
  std::vector<int> diag_off;  // [ndiag+1]
  int16_t * d_bricks = nullptr;
  uint32_t * d_plane_cells = nullptr;
  int32_t * d_plane_off = nullptr;
  void * d_fx = nullptr;
  void * d_fy = nullptr;
  void * d_fz = nullptr;
  size_t face_bytes = 0;
  int launches_per_sweep = 0;
  int autotuned = 0;

  // Per-launch HIP events (doma_profile_sweeps); recorded only while profiling.
  std::vector<hipEvent_t> prof_events;
  // Start/end events of the next sweep (null: untimed).  The persistent
  // schedules attach them to the kernel dispatch; the others record them around
  // their launches.
  hipEvent_t sweep_ev[2] = {nullptr, nullptr};
  bool profiling = false;

  // Dataflow schedule state.
  int schedule = 0;
  int ntx = 0, nty = 0, ntz = 0;
  int ntasks = 0;
  int grid_flow = 0;
  int16_t * d_pencils = nullptr;  // pencil schedule: [npencils][2]
  int npencils = 0;
  int nqueues = 1;  // ticket queues of the dataflow schedule
  uint32_t * d_flags = nullptr;
  unsigned long long * d_ticket = nullptr;
  int * d_err = nullptr;
  // Device-side source iteration: [0] done flag, [1] iterations (see IterArgs).
  int * d_si_state = nullptr;
  bool si_device = false;      // sweeps launched now check d_si_state
  bool si_chunked = true;      // source iteration with the stopping rule on the device
  std::vector<hipEvent_t> si_events;
  uint32_t epoch = 0;
  unsigned long long ticket_base = 0;
  int wall_clock_khz = 100000;

  // Learned SI correction (3D).
  doma::net::Net net;
  double * d_corr = nullptr;   // [N]
  double * d_dlast = nullptr;  // [N] last accepted increment
  double * d_dcur = nullptr;   // [N] increment of the latest sweep
  double * d_keep = nullptr;   // [N] flux before the pending correction
  int warm_start = 0;          // start group iterations from the previous MG iterate
  // Stopping rule of source iteration (DOMA_STOP_*) and, for the error-estimate
  // rule, the latest contraction estimate per group (kept across multigroup passes).
  int stop_rule = DOMA_STOP_RESIDUAL;
  std::vector<double> rho_est;
};

namespace doma
{
int create_ctx(const doma_problem_desc * d, hipStream_t shared, doma_ctx ** out);
int select_shape(int dims, int noct, int dirs, int minext, const doma_tuning & t);
void shape_chunking(int shape, int dirs, int * chunk, int * nchunks);
int configure_sweep(doma_ctx * c);
int run_sweep(doma_ctx * c, int g);
BrickArgs brick_args(doma_ctx * c, const double * q);
int check_sweep_error(doma_ctx * c);
int reset_after_sweep_error(doma_ctx * c);
void free_sweep(doma_ctx * c);
doma::IterArgs iter_args(doma_ctx * c, int g);
} // namespace doma
