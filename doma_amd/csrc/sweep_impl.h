// sweep_impl.h -- brick-wavefront transport sweep for CDNA4 (gfx950): kernel
// bodies.  Instantiated by sweep_*.hip; launched through sweep.hip.
//
// Decomposition (2D and 3D):
//   * The grid is cut into B^3 bricks (B^2 tiles in 2D).  Every octant visits
//     the bricks in its own sweep orientation; in oriented coordinates a brick
//     depends on its upwind neighbours along x, y and z only.  Vacuum boundaries
//     make the octants independent (TransportSolver3D.cpp:565-607,
//     CartesianCell3D.cpp:31-38), so all octants are swept at once.
//   * A work-group owns one oriented brick for KO octants and one direction
//     chunk at a time.  The KO octants see the same brick-local plane structure
//     (li+lj+lk = p in oriented coordinates), so their cells are packed into the
//     same passes; cells of a plane are independent and run concurrently.
//   * Threads are (octant, cell slot, direction lane).  DL consecutive lanes
//     carry NDL directions each, so the per-cell source and cross section are
//     uniform per lane group and a lane has NDL independent cell solves per pass.
//   * Everything a brick needs is staged in LDS before its first plane: upwind
//     face fluxes (B^2 x-, y- and z-faces per direction, updated in place), the
//     cell source and cross-section products, the global cell index and the
//     widths.  The plane loop touches no global memory except one store of the
//     direction-reduced scalar flux per cell; only the brick's inflow/outflow
//     faces cross HBM, through per-direction-set face planes.
//   * The scalar flux of a cell is reduced over its DL lanes with a fixed DPP
//     butterfly and stored once per (cell, direction set); the direction sets
//     are summed in a fixed order afterwards, so results are reproducible run to
//     run (no floating-point atomics).
//
// Three schedules share the brick body:
//   * diagonal: one launch per brick diagonal bi+bj+bk = s;
//   * dataflow: one persistent launch per sweep.  Work-groups take (brick,
//     octant group, chunk) tasks from monotone ticket counters in diagonal
//     order and wait on the epoch flags of the (at most three) upwind tasks.
//     The direction-set pairs are split over up to 16 counters (queues), each
//     served by its own share of the work-groups, so that no single counter
//     serialises the whole launch's atomics.  A task only waits on smaller
//     tickets of its own queue, taken by work-groups that are already running,
//     so the schedule cannot deadlock; every wait is bounded by a wall-clock
//     deadline and reports through an error flag.  Face planes are written and
//     read with device-coherent accesses, so no cache-wide write-back or
//     invalidate is needed per task;
//   * pencil (3D): the dataflow launch with a task per column of bricks along
//     z, swept by one work-group that keeps the z faces in LDS.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "cellsolve.h"
#include "internal.h"

namespace doma
{
namespace sweepk
{
// Kernel shapes: brick edge, direction lanes per cell, directions per lane,
// octants per work-group, threads per work-group.
struct Shape
{
  int dims, B, DL, NDL, KO, NT;
  int w32 = 0;  // FP32 kernels: waves per SIMD to budget registers for (0: compiler's choice)
};

// Shape table (index = shape id).  Chosen for direction chunks of <= 16, <= 32
// and <= 64 per octant, small and large bricks, in 3D and 2D.  A shape whose
// direction width DL * NDL is below the directions per octant splits them into
// several direction sets (more independent work-groups per brick diagonal).
constexpr Shape kShapes[] = {
    {3, 4, 8, 2, 4, 128},   // 0: <=16 dirs, 4^3 bricks, 4 octants
    {3, 4, 8, 2, 8, 256},   // 1: <=16 dirs, 4^3 bricks, 8 octants
    {3, 8, 8, 2, 2, 256},   // 2: <=16 dirs, 8^3 bricks, 2 octants
    {3, 4, 16, 2, 4, 256},  // 3: <=32 dirs, 4^3 bricks, 4 octants
    {3, 8, 16, 2, 1, 256},  // 4: <=32 dirs, 8^3 bricks
    {3, 8, 16, 2, 2, 512},  // 5: <=32 dirs, 8^3 bricks, 2 octants
    {3, 4, 32, 2, 2, 256},  // 6: <=64 dirs, 4^3 bricks, 2 octants
    {3, 8, 32, 2, 1, 256},  // 7: <=64 dirs, 8^3 bricks
    {3, 4, 8, 1, 4, 128},   // 8: <=8 dirs, 4^3 bricks, 4 octants
    {3, 8, 32, 1, 1, 512},  // 9: <=32 dirs, one direction per lane
    {2, 8, 8, 2, 4, 128},   // 10: 2D <=16 dirs, 8^2 tiles, 4 quadrants
    {2, 16, 16, 2, 2, 256}, // 11: 2D <=32 dirs, 16^2 tiles, 2 quadrants
    {2, 8, 32, 2, 1, 256},  // 12: 2D <=64 dirs, 8^2 tiles
    {2, 8, 8, 1, 4, 128},   // 13: 2D <=8 dirs
    {3, 4, 16, 1, 1, 256},  // 14: <=16 dirs, 4^3 bricks, one direction per lane
    {3, 4, 16, 1, 1, 128},  // 15: <=16 dirs, 4^3 bricks, 8 cells per pass
    {3, 8, 16, 1, 1, 512},  // 16: <=16 dirs, 8^3 bricks
    {3, 4, 32, 1, 1, 256},  // 17: <=32 dirs, 4^3 bricks
    {3, 8, 64, 1, 1, 512},  // 18: <=64 dirs, 8^3 bricks
    {2, 8, 16, 1, 1, 256},  // 19: 2D <=16 dirs
    {2, 16, 32, 1, 1, 512}, // 20: 2D <=32 dirs, 16^2 tiles
    {3, 8, 8, 1, 1, 512},   // 21: 8-direction sets, 8^3 bricks, 64 cells per pass
    {3, 8, 16, 1, 1, 1024}, // 22: <=16 dirs, 8^3 bricks, 64 cells per pass
    {3, 8, 8, 1, 1, 256},   // 23: 8-direction sets, 8^3 bricks, 32 cells per pass
    {3, 4, 8, 1, 1, 128},   // 24: 8-direction sets, 4^3 bricks
    {3, 8, 4, 1, 1, 256},   // 25: 4-direction sets, 8^3 bricks, 64 cells per pass
    {3, 8, 8, 1, 2, 512},   // 26: 8-direction sets, 2 octants share a pass (32 cells each)
    {3, 8, 4, 1, 4, 512},   // 27: 4-direction sets, 4 octants share a pass (32 cells each)
    {3, 4, 4, 1, 8, 256},   // 28: 4-direction sets, 4^3 bricks, 8 octants share a pass
    {3, 4, 4, 1, 1, 64},    // 29: 4-direction sets, 4^3 bricks, one wave
    {3, 4, 4, 1, 2, 128},   // 30: 4-direction sets, 4^3 bricks, 2 octants
    {3, 4, 8, 1, 1, 128},   // 31: 8-direction sets, 4^3 bricks, 16 cells per pass
    {3, 8, 8, 4, 1, 256, 4}, // 32: <=32 dirs, 8^3 bricks, four directions per lane, 32 cells per pass
};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// ------------------------------------------------------------------------- //
// Direction reduction                                                        //
// ------------------------------------------------------------------------- //

template <int CTRL>
__device__ __forceinline__ double
dpp_mov(double v)
{
  // Every lane of the permutations used here reads a valid lane of its own row,
  // so no lane takes the bound-control value and no old value is needed.
  const int lo = __double2loint(v);
  const int hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
  const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
  return __hiloint2double(rhi, rlo);
}

// DPP controls (GFX9 encoding): quad_perm(1,0,3,2), quad_perm(2,3,0,1),
// row_half_mirror, row_mirror.
constexpr int DPP_QP_XOR1 = 0xB1;
constexpr int DPP_QP_XOR2 = 0x4E;
constexpr int DPP_HALF_MIRROR = 0x141;
constexpr int DPP_MIRROR = 0x140;

// Row broadcasts (GFX9 DPP): lane 15 of rows 0 and 2 to rows 1 and 3
// (row_mask 0xA), lane 31 to rows 2 and 3 (row_mask 0xC).  Rows outside the
// mask are not written and hold an unspecified value.
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double
dpp_bcast(double v)
{
  const int lo = __double2loint(v);
  const int hi = __double2hiint(v);
  const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, ROW_MASK, 0xF, false);
  const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, ROW_MASK, 0xF, false);
  return __hiloint2double(rhi, rlo);
}

// Fixed-order sum over groups of DL (4..64) consecutive lanes.  The sum is
// complete in the group's last lane (for DL <= 16 in every lane of the group);
// each step adds a commutative pair, so the result is the same bitwise whichever
// lane of a pair receives it.  Across the 16-lane rows the row sums move by
// row broadcasts rather than a cross-lane permute through LDS.
template <int DL>
__device__ __forceinline__ double
group_sum(double v)
{
  static_assert(DL >= 4 && DL <= 64, "direction lanes");
  v += dpp_mov<DPP_QP_XOR1>(v);
  v += dpp_mov<DPP_QP_XOR2>(v);
  if constexpr (DL >= 8)
    v += dpp_mov<DPP_HALF_MIRROR>(v);
  if constexpr (DL >= 16)
    v += dpp_mov<DPP_MIRROR>(v);
  if constexpr (DL >= 32)
    v += dpp_bcast<DPP_ROW_BCAST15, 0xA>(v);  // rows 1, 3: r0 + r1, r2 + r3
  if constexpr (DL >= 64)
    v += dpp_bcast<DPP_ROW_BCAST31, 0xC>(v);  // row 3: (r2 + r3) + (r0 + r1)
  return v;
}

// ------------------------------------------------------------------------- //
// LDS layout                                                                 //
// ------------------------------------------------------------------------- //

// Per-cell record staged in LDS for a brick: source and cross-section terms,
// the geometry factors the cell solve needs, and the global cell index (-1 for
// a cell of a partial brick outside the grid).
//   fast TW-DD 3D: v0 = q V, v1 = sigma V, g = (ly lz, lx lz, lx ly)
//   fast TW-DD 2D: v0 = q V, v1 = sigma V, g = (ly, lx, 1)
//   fast DD:       v0 = q,   v1 = sigma,   g = (1/lx, 1/ly, 1/lz)
//   exact:         v0 = q,   v1 = sigma,   g = (lx, ly, lz)
// The fields have the sweep's arithmetic type (converted once, when the record
// is built): 48 bytes in FP64, 24 in FP32.
template <typename R>
struct alignas(sizeof(R) * 2) CellRec
{
  R v0, v1, g0, g1, g2;
  int c;
};
static_assert(sizeof(CellRec<double>) == 48 && sizeof(CellRec<float>) == 24, "cell records");

// LDS bytes of a brick layout with records of `rec` bytes and plane-table
// entries of `ptb` bytes.
constexpr size_t
layout_bytes(size_t face_bytes, int nc, int nplanes, int ko, size_t ptb, size_t rec)
{
  return ((face_bytes + 15) / 16 * 16 + rec * (size_t)nc * ko + ptb * nc + 4 * (size_t)(nplanes + 1) +
          2 * (size_t)nc + 15) / 16 * 16 + 16;
}

// Work-groups of `bytes` of LDS each that a CU's 160 KiB holds (allocated in
// 512-byte granules).
constexpr size_t
lds_groups(size_t bytes)
{
  return (160 * 1024) / ((bytes + 511) / 512 * 512);
}

// Plane-table entries are 8 bytes (WIDE) unless that lowers the work-groups a
// CU's LDS holds; then 4.
constexpr bool
wide_plane_table(size_t face_bytes, int nc, int nplanes, int ko, size_t rec)
{
  return lds_groups(layout_bytes(face_bytes, nc, nplanes, ko, 8, rec)) ==
         lds_groups(layout_bytes(face_bytes, nc, nplanes, ko, 4, rec));
}

template <typename R, int DIM, int B, int DL, int NDL, int KO>
struct Layout
{
  static constexpr int DW = DL * NDL;                      // direction slots per face slot
  static constexpr int FS = DIM == 3 ? B * B : B;          // face slots per axis
  static constexpr int NC = DIM == 3 ? B * B * B : B * B;  // cells per brick
  static constexpr int NPLANES = DIM == 3 ? 3 * B - 2 : 2 * B - 1;
  static constexpr int NAX = DIM == 3 ? 3 : 2;
  static constexpr size_t FACE = sizeof(R) * FS * DW;      // one axis, one octant
  static constexpr size_t REC = sizeof(CellRec<R>);
  static constexpr bool WIDE = wide_plane_table(FACE * NAX * KO, NC, NPLANES, KO, REC);
  static constexpr size_t PTB = WIDE ? 8 : 4;              // plane-table entry bytes
  static constexpr size_t FACES_OFF = 0;
  static constexpr size_t REC_OFF = ((FACE * NAX * KO) + 15) / 16 * 16;   // CellRec [KO][NC]
  static constexpr size_t PT_OFF = REC_OFF + REC * NC * KO;               // entries [NC]
  static constexpr size_t PO_OFF = PT_OFF + PTB * NC;                      // int [NPLANES+1]
  static constexpr size_t PI_OFF = PO_OFF + 4 * (NPLANES + 1);                // uint16 [NC]
  static constexpr size_t TASK_OFF = (PI_OFF + 2 * NC + 15) / 16 * 16;
  static constexpr size_t BYTES = TASK_OFF + 16;
  static_assert(BYTES == layout_bytes(FACE * NAX * KO, NC, NPLANES, KO, PTB, REC), "LDS layout");
};

// Register budget hint of a kernel: the shape's w32 waves per SIMD for FP32
// sweeps whose work-group LDS leaves room for that many (so that registers do
// not hold the work-groups per CU below what LDS allows), else none.  Shape 32
// (the FP32 DD sweep's tuned shape) fits 4 work-groups of 4 waves per CU in LDS
// and needs 131 VGPRs unhinted, 127 at a 4-wave budget, without spills.
template <typename R, int DIM, int SHAPE>
constexpr int
waves_per_simd()
{
  constexpr Shape S = kShapes[SHAPE];
  using L = Layout<R, DIM, S.B, S.DL, S.NDL, S.KO>;
  constexpr int w = (int)lds_groups(L::BYTES) * (S.NT / 64) / 4;
  return sizeof(R) == 4 && S.w32 > 1 && w >= S.w32 ? S.w32 : 1;
}

// Plane table: per position of the plane-ordered cell list, where the cell's x-,
// y- and z-face slots lie (3D: sx = lj B + lk, sy = li B + lk, sz = li B + lj;
// 2D: sx = lj, sy = li; slot s of an axis starts s DW sizeof(R) bytes into it)
// and the brick-local cell ce.  WIDE entries hold the byte offsets, as
// {x | y << 16, z | ce << 16}, so a lane finds its faces with one add per axis;
// narrow ones pack the slots as ce | sx << 9 | sy << 15 | sz << 21.
typedef uint32_t PlaneEntry __attribute__((ext_vector_type(2)));

template <bool WIDE>
using PlaneEntryT = std::conditional_t<WIDE, PlaneEntry, uint32_t>;

template <typename R, int DIM, int B, int DL, int NDL, int KO, int NT>
__device__ __forceinline__ void
load_plane_table(const BrickArgs & a, unsigned char * smem)
{
  using L = Layout<R, DIM, B, DL, NDL, KO>;
  static_assert(L::NC <= 512 && L::FS <= 64 && (L::FS - 1) * L::DW * sizeof(R) < 65536,
                "plane table packing");
  PlaneEntryT<L::WIDE> * pt = reinterpret_cast<PlaneEntryT<L::WIDE> *>(smem + L::PT_OFF);
  int32_t * po = reinterpret_cast<int32_t *>(smem + L::PO_OFF);
  uint16_t * pi = reinterpret_cast<uint16_t *>(smem + L::PI_OFF);
  for (int e = threadIdx.x; e < L::NC; e += NT)
  {
    const uint32_t ce = a.plane_cells[e];
    pi[ce] = (uint16_t)e;
    const uint32_t li = ce % B, lj = (ce / B) % B, lk = DIM == 3 ? ce / (B * B) : 0;
    const uint32_t sx = DIM == 3 ? lj * B + lk : lj;
    const uint32_t sy = DIM == 3 ? li * B + lk : li;
    const uint32_t sz = DIM == 3 ? li * B + lj : 0;
    if constexpr (L::WIDE)
    {
      constexpr uint32_t step = L::DW * sizeof(R);
      pt[e] = PlaneEntry{sx * step | (sy * step) << 16, sz * step | ce << 16};
    }
    else
      pt[e] = ce | (sx << 9) | (sy << 15) | (sz << 21);
  }
  for (int e = threadIdx.x; e <= L::NPLANES; e += NT)
    po[e] = a.plane_off[e];
}

// Orientation of an octant (3D: 4=-mu 2=-eta 1=-xi, TransportBase.h:22-32) or a
// 2D quadrant slot (PPP, PMP, MPP, MMP: 2=-mu 1=-eta).
template <int DIM>
__device__ __forceinline__ void
orient(int oct, bool & nx, bool & ny, bool & nz)
{
  nx = DIM == 3 ? (oct & 4) : (oct & 2);
  ny = DIM == 3 ? (oct & 2) : (oct & 1);
  nz = DIM == 3 ? (oct & 1) : false;
}

// Device-coherent face-plane accesses for the dataflow schedule (bypass the
// per-XCD L2 so that a consumer on another XCD sees the producer's stores once
// they have completed); plain accesses for the diagonal schedule, where the
// kernel boundary orders producer and consumer.
template <bool COHERENT, typename R>
__device__ __forceinline__ R
face_load(const R * p)
{
  if constexpr (COHERENT)
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    return *p;
}

template <bool COHERENT, typename R>
__device__ __forceinline__ void
face_store(R * p, R v)
{
  if constexpr (COHERENT)
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    *p = v;
}

// ------------------------------------------------------------------------- //
// Brick body                                                                 //
// ------------------------------------------------------------------------- //

// No upwind dependency to wait for (the diagonal schedule orders bricks by launch).
struct NoWait
{
  __device__ void operator()() const {}
};

// PENCIL: the brick is one of a column along z swept by the same work-group;
// its z faces never cross global memory (zero inflow at bk = 0, else the
// previous brick's outflow, still in LDS).
template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NDL, int KO, int NT,
          bool COHERENT, bool PENCIL = false, typename Wait = NoWait>
__device__ __forceinline__ void
sweep_brick(const BrickArgs & a, int bi, int bj, int bk, int og, int ch, unsigned char * smem,
            Wait wait_upwind = Wait{})
{
  using L = Layout<R, DIM, B, DL, NDL, KO>;
  constexpr int DW = L::DW, FS = L::FS, NC = L::NC, NAX = L::NAX;
  constexpr int CPP = NT / DL;   // cell slots per pass
  constexpr int CPO = CPP / KO;  // cells per octant per pass
  static_assert(CPP % KO == 0, "cell slots must split evenly over octants");

  R * faces = reinterpret_cast<R *>(smem + L::FACES_OFF);  // [KO][NAX][FS][DW]
  CellRec<R> * rec = reinterpret_cast<CellRec<R> *>(smem + L::REC_OFF);
  using PE = PlaneEntryT<L::WIDE>;
  const PE * pt = reinterpret_cast<const PE *>(smem + L::PT_OFF);
  const int32_t * po = reinterpret_cast<const int32_t *>(smem + L::PO_OFF);
  const uint16_t * pi = reinterpret_cast<const uint16_t *>(smem + L::PI_OFF);

  const int tid = threadIdx.x;
  const int dd = tid % DL;
  const int slot = tid / DL;
  const int ko = slot % KO;
  const int cpo = slot / KO;
  const int oct = og * KO + ko;
  const int ds = oct * a.nchunks + ch;

  R amu[NDL], aeta[NDL], axi[NDL];
  double w[NDL];
#pragma unroll
  for (int m = 0; m < NDL; ++m)
  {
    const int lane_dir = dd * NDL + m;
    const int n = ch * a.chunk + lane_dir;
    amu[m] = aeta[m] = axi[m] = R(1);
    w[m] = 0.0;
    if (lane_dir < a.chunk && n < a.dirs)
    {
      const double * t = a.dirtab + 4 * (oct * a.dirs + n);
      amu[m] = R(t[0]);
      aeta[m] = R(t[1]);
      axi[m] = R(t[2]);
      w[m] = t[3];
    }
  }

  const int i0 = bi * B, j0 = bj * B, k0 = bk * B;
  R * gf[3] = {reinterpret_cast<R *>(a.fx), reinterpret_cast<R *>(a.fy),
               reinterpret_cast<R *>(a.fz)};

  // The previous task of this work-group may still be reading LDS.
  __syncthreads();

  // Cell records of every octant, in plane-table order (cells are read in
  // brick-local order ce = (lk*B + lj)*B + li and stored at position pi[ce]), and the
  // inflow faces of every octant (the upwind brick's outflow, or vacuum).  Each
  // lane issues its loads in groups before waiting for any: the first group of
  // cell inputs (which no upwind task produces) before the wait for the upwind
  // tasks, then the faces, so one memory latency covers both.
  constexpr int NR = KO * NC;             // records
  constexpr int RI = (NR + NT - 1) / NT;  // per lane
  constexpr int RG = RI < 4 ? RI : 4;     // cells whose inputs are in flight per group
  auto load_cells = [&](int r0, double (&q)[RG], double (&sg)[RG], double (&lx)[RG],
                        double (&ly)[RG], double (&lz)[RG], int (&cell)[RG])
                        __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < RG; ++k)
    {
      const int e = tid + (r0 + k) * NT;
      cell[k] = -1;
      q[k] = 0.0;
      sg[k] = lx[k] = ly[k] = lz[k] = 1.0;
      if (r0 + k < RI && e < NR)
      {
        const int k_o = e / NC, ce = e - k_o * NC;
        bool nx, ny, nz;
        orient<DIM>(og * KO + k_o, nx, ny, nz);
        const int li = ce % B, lj = (ce / B) % B, lk = DIM == 3 ? ce / (B * B) : 0;
        const int oi = i0 + li, oj = j0 + lj, ok = k0 + lk;
        if (oi < a.nx && oj < a.ny && ok < a.nz)
        {
          const int i = nx ? a.nx - 1 - oi : oi;
          const int j = ny ? a.ny - 1 - oj : oj;
          const int k3 = nz ? a.nz - 1 - ok : ok;
          const int c = (k3 * a.ny + j) * a.nx + i;
          cell[k] = c;
          q[k] = a.q[c];
          sg[k] = a.sig[c];
          lx[k] = a.lx[i];
          ly[k] = a.ly[j];
          lz[k] = DIM == 3 ? a.lz[k3] : 1.0;
        }
      }
    }
  };
  auto store_cells = [&](int r0, const double (&q)[RG], const double (&sg)[RG],
                         const double (&lx)[RG], const double (&ly)[RG], const double (&lz)[RG],
                         const int (&cell)[RG]) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < RG; ++k)
    {
      const int e = tid + (r0 + k) * NT;
      if (!(r0 + k < RI && e < NR))
        continue;
      // The terms are formed in FP64 and rounded to R once, here.
      double v0 = 0.0, v1 = 1.0, g0 = 1.0, g1 = 1.0, g2 = 1.0;
      if (cell[k] >= 0)
      {
        if (EXACT)
        {
          v0 = q[k];
          v1 = sg[k];
          g0 = lx[k];
          g1 = ly[k];
          g2 = lz[k];
        }
        else if (SCHEME == DOMA_SCHEME_DD)
        {
          v0 = q[k];
          v1 = sg[k];
          g0 = 1.0 / lx[k];
          g1 = 1.0 / ly[k];
          g2 = 1.0 / lz[k];
        }
        else
        {
          const double V = DIM == 3 ? lx[k] * ly[k] * lz[k] : lx[k] * ly[k];
          v0 = q[k] * V;
          v1 = sg[k] * V;
          g0 = DIM == 3 ? ly[k] * lz[k] : ly[k];
          g1 = DIM == 3 ? lx[k] * lz[k] : lx[k];
          g2 = DIM == 3 ? lx[k] * ly[k] : 1.0;
        }
      }
      CellRec<R> r;
      r.v0 = R(v0);
      r.v1 = R(v1);
      r.g0 = R(g0);
      r.g1 = R(g1);
      r.g2 = R(g2);
      r.c = cell[k];
      const int k_o = e / NC;
      rec[k_o * NC + pi[e - k_o * NC]] = r;
    }
  };
  double cq[RG], csg[RG], clx[RG], cly[RG], clz[RG];
  int ccell[RG];
  load_cells(0, cq, csg, clx, cly, clz, ccell);

  // Upwind tasks of this sweep have published their outflow faces.
  wait_upwind();
  __syncthreads();

  static_assert(!PENCIL || DIM == 3, "pencils are columns of 3D bricks");
  constexpr int NAXG = PENCIL ? 2 : NAX;   // axes whose faces cross global memory
  constexpr int NF = KO * NAXG * FS * DW;  // face values loaded
  constexpr int FI = (NF + NT - 1) / NT;   // per lane
  // LDS index of loaded face value e (the LDS keeps all NAX axes per octant).
  auto face_slot = [](int e) {
    return PENCIL ? (e / (NAXG * FS * DW)) * (NAX * FS * DW) + e % (NAXG * FS * DW) : e;
  };
  if constexpr (PENCIL)
  {
    if (bk == 0)
      for (int e = tid; e < KO * FS * DW; e += NT)
        faces[(e / (FS * DW)) * (NAX * FS * DW) + 2 * FS * DW + e % (FS * DW)] = R(0);
  }
  constexpr int FG = FI < 4 ? FI : 4;      // loads in flight per group
#pragma unroll
  for (int f0 = 0; f0 < FI; f0 += FG)
  {
    R val[FG];
#pragma unroll
    for (int k = 0; k < FG; ++k)
    {
      const int e = tid + (f0 + k) * NT;
      val[k] = R(0);
      if (f0 + k < FI && e < NF)
      {
        const int k_o = e / (NAXG * FS * DW);
        const int r = e - k_o * (NAXG * FS * DW);
        const int ax = r / (FS * DW);
        const int s = (r / DW) % FS, l = r % DW;
        const int u = DIM == 3 ? s / B : s;
        const int v = DIM == 3 ? s % B : 0;
        const int dsk = (og * KO + k_o) * a.nchunks + ch;
        if (ax == 0 && bi > 0)
          val[k] = face_load<COHERENT>(gf[0] + (DIM == 3 ? ((size_t)dsk * a.nyp * a.nzp +
                                                            (size_t)(j0 + u) * a.nzp + (k0 + v))
                                                         : ((size_t)dsk * a.nyp + (j0 + u))) * DW + l);
        else if (ax == 1 && bj > 0)
          val[k] = face_load<COHERENT>(gf[1] + (DIM == 3 ? ((size_t)dsk * a.nxp * a.nzp +
                                                            (size_t)(i0 + u) * a.nzp + (k0 + v))
                                                         : ((size_t)dsk * a.nxp + (i0 + u))) * DW + l);
        else if (ax == 2 && bk > 0)
          val[k] = face_load<COHERENT>(gf[2] + ((size_t)dsk * a.nxp * a.nyp +
                                                (size_t)(i0 + u) * a.nyp + (j0 + v)) * DW + l);
      }
    }
#pragma unroll
    for (int k = 0; k < FG; ++k)
    {
      const int e = tid + (f0 + k) * NT;
      if (f0 + k < FI && e < NF)
        faces[face_slot(e)] = val[k];
    }
  }
  store_cells(0, cq, csg, clx, cly, clz, ccell);
#pragma unroll
  for (int r0 = RG; r0 < RI; r0 += RG)
  {
    load_cells(r0, cq, csg, clx, cly, clz, ccell);
    store_cells(r0, cq, csg, clx, cly, clz, ccell);
  }
  __syncthreads();

  R * fxo = faces + (size_t)ko * NAX * FS * DW + dd * NDL;  // this lane's octant and directions
  R * fyo = fxo + FS * DW;
  R * fzo = fyo + FS * DW;
  const CellRec<R> * rco = rec + ko * NC;

  // One cell solve of the lane: x-, y- and z-faces in LDS updated in place, the
  // direction-reduced scalar flux stored once per (cell, direction set).
  auto solve = [&](PE info, const CellRec<R> & r) __attribute__((always_inline)) {
    R * xf, *yf, *zf;
    if constexpr (L::WIDE)
    {
      xf = reinterpret_cast<R *>(reinterpret_cast<unsigned char *>(fxo) + (info.x & 0xFFFFu));
      yf = reinterpret_cast<R *>(reinterpret_cast<unsigned char *>(fyo) + (info.x >> 16));
      zf = reinterpret_cast<R *>(reinterpret_cast<unsigned char *>(fzo) + (info.y & 0xFFFFu));
    }
    else
    {
      xf = fxo + ((info >> 9) & 63u) * DW;
      yf = fyo + ((info >> 15) & 63u) * DW;
      zf = fzo + ((info >> 21) & 63u) * DW;
    }
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < NDL; ++m)
    {
      R x = xf[m], y = yf[m];
      R z = DIM == 3 ? zf[m] : R(0);
      R psi;
      if constexpr (DIM == 3)
      {
        if constexpr (SCHEME == DOMA_SCHEME_TW_DD)
        {
          if constexpr (EXACT)
            psi = twdd3d_exact<R>(R(r.v0), R(r.v1), R(r.g0), R(r.g1), R(r.g2), amu[m], aeta[m],
                                  axi[m], x, y, z);
          else
            psi = twdd3d_fast<R>(R(r.v0), R(r.v1), amu[m] * R(r.g0), aeta[m] * R(r.g1),
                                 axi[m] * R(r.g2), x, y, z);
        }
        else
        {
          if constexpr (EXACT)
            psi = dd3d_exact<R>(R(r.v0), R(r.v1), R(r.g0), R(r.g1), R(r.g2), amu[m], aeta[m],
                                axi[m], x, y, z);
          else
            psi = dd3d_fast<R>(R(r.v0), R(r.v1), R(2) * amu[m] * R(r.g0),
                               R(2) * aeta[m] * R(r.g1), R(2) * axi[m] * R(r.g2), x, y, z);
        }
      }
      else
      {
        if constexpr (SCHEME == DOMA_SCHEME_TW_DD)
        {
          if constexpr (EXACT)
            psi = twdd2d_exact<R>(R(r.v0), R(r.v1), R(r.g0), R(r.g1), amu[m], aeta[m], axi[m], x,
                                  y);
          else
            psi = twdd2d_fast<R>(R(r.v0), R(r.v1), amu[m] * R(r.g0), aeta[m] * R(r.g1), axi[m], x,
                                 y);
        }
        else
        {
          if constexpr (EXACT)
            psi = dd2d_exact<R>(R(r.v0), R(r.v1), R(r.g0), R(r.g1), amu[m], aeta[m], x, y);
          else
            psi = dd3d_fast<R>(R(r.v0), R(r.v1), R(2) * amu[m] * R(r.g0),
                               R(2) * aeta[m] * R(r.g1), R(0), x, y, z);
        }
      }
      xf[m] = x;
      yf[m] = y;
      if constexpr (DIM == 3)
        zf[m] = z;
      s += w[m] * static_cast<double>(psi);
    }
    s = group_sum<DL>(s);
    if (dd == DL - 1 && r.c >= 0)
      a.part[(size_t)ds * a.N + r.c] = s;
  };

  // The lane's first cell of each plane: its plane-table entry and record are
  // read at the end of the previous plane, before its barrier (neither changes
  // during the task), so after the barrier only the face reads stand between
  // the lane and its next solve.  Records sit in plane-table order, so the
  // record read does not wait for the plane-table entry.  Cells of a partial
  // brick outside the grid are solved too (on finite dummy data): they only
  // feed other outside cells, and their scalar flux is not stored.
  int pb = po[0] + cpo, pe = po[1];
  PE info_n{};
  CellRec<R> r_n{};
  if (pb < pe)
  {
    info_n = pt[pb];
    r_n = rco[pb];
  }
  for (int p = 0; p < L::NPLANES; ++p)
  {
    const int nb = pe + cpo, ne = p + 1 < L::NPLANES ? po[p + 2] : pe;
    if (pb < pe)
    {
      solve(info_n, r_n);
      for (int pos = pb + CPO; pos < pe; pos += CPO)
        solve(pt[pos], rco[pos]);
    }
    if (p + 1 < L::NPLANES && nb < ne)
    {
      info_n = pt[nb];
      r_n = rco[nb];
    }
    pb = nb;
    pe = ne;
    __syncthreads();
  }

  // Outflow faces for the downwind bricks (padded extents keep every index in range).
  for (int e = tid; e < NF; e += NT)
  {
    const int k_o = e / (NAXG * FS * DW);
    const int r = e - k_o * (NAXG * FS * DW);
    const int ax = r / (FS * DW);
    const int s = (r / DW) % FS, l = r % DW;
    const int u = DIM == 3 ? s / B : s;
    const int v = DIM == 3 ? s % B : 0;
    const int dsk = (og * KO + k_o) * a.nchunks + ch;
    R * dst;
    if (ax == 0)
      dst = gf[0] + (DIM == 3 ? ((size_t)dsk * a.nyp * a.nzp + (size_t)(j0 + u) * a.nzp + (k0 + v))
                              : ((size_t)dsk * a.nyp + (j0 + u))) * DW + l;
    else if (ax == 1)
      dst = gf[1] + (DIM == 3 ? ((size_t)dsk * a.nxp * a.nzp + (size_t)(i0 + u) * a.nzp + (k0 + v))
                              : ((size_t)dsk * a.nxp + (i0 + u))) * DW + l;
    else
      dst = gf[2] + ((size_t)dsk * a.nxp * a.nyp + (size_t)(i0 + u) * a.nyp + (j0 + v)) * DW + l;
    face_store<COHERENT>(dst, faces[face_slot(e)]);
  }
}

// One launch per brick diagonal: block = (brick on the diagonal, octant group, chunk).
template <typename R, int DIM, int SCHEME, bool EXACT, int SHAPE>
__global__ __launch_bounds__(kShapes[SHAPE].NT)
__attribute__((amdgpu_waves_per_eu(waves_per_simd<R, DIM, SHAPE>()))) void
brick_sweep_kernel(BrickArgs a)
{
  if (a.skip && *a.skip)
    return;
  constexpr Shape S = kShapes[SHAPE];
  extern __shared__ __align__(16) unsigned char smem[];
  load_plane_table<R, DIM, S.B, S.DL, S.NDL, S.KO, S.NT>(a, smem);
  const int ngroups = a.ndirsets / (a.nchunks * S.KO);  // octant groups
  const int per_brick = ngroups * a.nchunks;
  const int br = a.diag_begin + blockIdx.x / per_brick;
  const int r = blockIdx.x % per_brick;
  const int og = r / a.nchunks, ch = r % a.nchunks;
  sweep_brick<R, DIM, SCHEME, EXACT, S.B, S.DL, S.NDL, S.KO, S.NT, false>(
      a, a.bricks[3 * br + 0], a.bricks[3 * br + 1], a.bricks[3 * br + 2], og, ch, smem);
}

// Bounded wait for an upwind task's epoch flag (thread 0 only).  Polls with
// relaxed device-coherent loads; gives up at a wall-clock deadline, or at once
// when another wait of this sweep has already expired.
//
// Visibility of the upwind faces does not rest on release/acquire semantics:
// the producer's face stores are device-coherent (they bypass the per-XCD L2)
// and have completed (s_waitcnt) before its flag store, and the consumer's face
// loads are device-coherent too.  The signal fences below only stop the
// compiler from moving face accesses across the flag accesses; they emit no
// instruction.
__device__ __forceinline__ void
wait_epoch_body(const uint32_t * f, uint32_t epoch, int * error, long long wait_ticks)
{
  if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= epoch)
    return;
  const long long t0 = wall_clock64();
  int polls = 0;
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch)
  {
    __builtin_amdgcn_s_sleep(2);
    if ((++polls & 31) == 0)
    {
      if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        return;
      if (wall_clock64() - t0 > wait_ticks)
      {
        atomicExch(error, 1);
        return;
      }
    }
  }
}

__device__ __forceinline__ void
wait_epoch(const uint32_t * f, uint32_t epoch, int * error, long long wait_ticks)
{
  wait_epoch_body(f, epoch, error, wait_ticks);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Publishes a finished task's epoch (thread 0, after every thread's face
// stores have completed and the work-group has synchronised).
__device__ __forceinline__ void
publish_epoch(uint32_t * f, uint32_t epoch)
{
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_store(f, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void
wait_flag(const BrickArgs & a, const uint32_t * f)
{
  wait_epoch(f, a.epoch, a.error, a.wait_ticks);
}

// One persistent launch per sweep; tasks in ticket (= diagonal) order, task id
// = (brick, octant group, chunk).
template <typename R, int DIM, int SCHEME, bool EXACT, int SHAPE>
__global__ __launch_bounds__(kShapes[SHAPE].NT)
__attribute__((amdgpu_waves_per_eu(waves_per_simd<R, DIM, SHAPE>()))) void
brick_flow_kernel(BrickArgs a)
{
  if (a.skip && *a.skip)
  {
    // Keep the ticket counters where a full launch would leave them.
    if (blockIdx.x == 0 && (int)threadIdx.x < a.nqueues)
      atomicAdd(a.ticket + (size_t)threadIdx.x * kTicketStride,
                ((unsigned long long)a.ntasks + gridDim.x) / a.nqueues);
    return;
  }
  constexpr Shape S = kShapes[SHAPE];
  using L = Layout<R, DIM, S.B, S.DL, S.NDL, S.KO>;
  extern __shared__ __align__(16) unsigned char smem[];
  int * s_task = reinterpret_cast<int *>(smem + L::TASK_OFF);
  load_plane_table<R, DIM, S.B, S.DL, S.NDL, S.KO, S.NT>(a, smem);
  const int nbricks = a.ntx * a.nty * a.ntz;
  const int per_brick = a.ndirsets / S.KO;  // (octant group, chunk) pairs
  // This work-group's ticket queue: pairs r = queue, queue + nq, ... of every
  // brick, in brick (= diagonal) order.  Upwind tasks share the pair and so the
  // queue, and hold smaller tickets in it.
  const int nq = a.nqueues;
  const int queue = (int)(blockIdx.x % nq);
  const int pq = per_brick / nq;  // pairs per brick in one queue
  unsigned long long * ticket = a.ticket + (size_t)queue * kTicketStride;
  const unsigned long long base = a.ticket_base / nq;
  const unsigned long long ntq = (unsigned long long)(a.ntasks / nq);
  while (true)
  {
    // A barrier heads every iteration so that the divergent "publish flag" of
    // one task and "take ticket" of the next can never be merged across the
    // back-edge into a nested loop that lane 0 leaves without its wave.
    __syncthreads();
    if (threadIdx.x == 0)
    {
      const unsigned long long u = atomicAdd(ticket, 1ull) - base;
      *s_task = u < ntq ? (int)u : -1;
    }
    __syncthreads();
    // Uniform by construction; readfirstlane lets the compiler see it.
    const int u = __builtin_amdgcn_readfirstlane(*s_task);
    if (u < 0)
      break;
    const int br = u / pq;
    const int r = (u - br * pq) * nq + queue;
    const int og = r / a.nchunks, ch = r % a.nchunks;
    const int bi = a.bricks[3 * br + 0], bj = a.bricks[3 * br + 1], bk = a.bricks[3 * br + 2];
    uint32_t * flags = a.flags + (size_t)r * nbricks;
    auto wait_upwind = [&]() {
      if (threadIdx.x == 0)
      {
        if (bi > 0)
          wait_flag(a, flags + (bk * a.nty + bj) * a.ntx + (bi - 1));
        if (bj > 0)
          wait_flag(a, flags + (bk * a.nty + (bj - 1)) * a.ntx + bi);
        if (bk > 0)
          wait_flag(a, flags + ((bk - 1) * a.nty + bj) * a.ntx + bi);
      }
    };
    sweep_brick<R, DIM, SCHEME, EXACT, S.B, S.DL, S.NDL, S.KO, S.NT, true>(a, bi, bj, bk, og, ch,
                                                                           smem, wait_upwind);
    // Every thread's coherent face stores must have completed before the flag.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
      publish_epoch(flags + (bk * a.nty + bj) * a.ntx + bi, a.epoch);
  }
}

// One persistent launch per sweep whose tasks are pencils: (bi, bj, all bk) for
// one (octant group, chunk) pair.  The work-group sweeps the column's bricks in
// z order and keeps their z faces in LDS, so only the x and y faces cross
// global memory and only the x and y neighbours are waited for.  Pencils are
// ticketed in bi+bj order per queue as bricks are in brick_flow_kernel; a
// pencil's brick bk waits only on brick bk of the two upwind pencils, which hold
// smaller tickets of the same queue, so the induction that rules out deadlock
// there holds brick by brick.
template <typename R, int DIM, int SCHEME, bool EXACT, int SHAPE>
__global__ __launch_bounds__(kShapes[SHAPE].NT)
__attribute__((amdgpu_waves_per_eu(waves_per_simd<R, DIM, SHAPE>()))) void
brick_pencil_kernel(BrickArgs a)
{
  if (a.skip && *a.skip)
  {
    if (blockIdx.x == 0 && (int)threadIdx.x < a.nqueues)
      atomicAdd(a.ticket + (size_t)threadIdx.x * kTicketStride,
                ((unsigned long long)a.ntasks + gridDim.x) / a.nqueues);
    return;
  }
  constexpr Shape S = kShapes[SHAPE];
  using L = Layout<R, DIM, S.B, S.DL, S.NDL, S.KO>;
  extern __shared__ __align__(16) unsigned char smem[];
  int * s_task = reinterpret_cast<int *>(smem + L::TASK_OFF);
  load_plane_table<R, DIM, S.B, S.DL, S.NDL, S.KO, S.NT>(a, smem);
  const int nbricks = a.ntx * a.nty * a.ntz;
  const int per_pencil = a.ndirsets / S.KO;  // (octant group, chunk) pairs
  const int nq = a.nqueues;
  const int queue = (int)(blockIdx.x % nq);
  const int pq = per_pencil / nq;
  unsigned long long * ticket = a.ticket + (size_t)queue * kTicketStride;
  const unsigned long long base = a.ticket_base / nq;
  const unsigned long long ntq = (unsigned long long)(a.ntasks / nq);
  while (true)
  {
    __syncthreads();  // see brick_flow_kernel
    if (threadIdx.x == 0)
    {
      const unsigned long long u = atomicAdd(ticket, 1ull) - base;
      *s_task = u < ntq ? (int)u : -1;
    }
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(*s_task);
    if (u < 0)
      break;
    const int pc = u / pq;
    const int r = (u - pc * pq) * nq + queue;
    const int og = r / a.nchunks, ch = r % a.nchunks;
    const int bi = a.pencils[2 * pc + 0], bj = a.pencils[2 * pc + 1];
    uint32_t * flags = a.flags + (size_t)r * nbricks;
    for (int bk = 0; bk < a.ntz; ++bk)
    {
      auto wait_upwind = [&]() {
        if (threadIdx.x == 0)
        {
          if (bi > 0)
            wait_flag(a, flags + (bk * a.nty + bj) * a.ntx + (bi - 1));
          if (bj > 0)
            wait_flag(a, flags + (bk * a.nty + (bj - 1)) * a.ntx + bi);
        }
      };
      sweep_brick<R, DIM, SCHEME, EXACT, S.B, S.DL, S.NDL, S.KO, S.NT, true, true>(
          a, bi, bj, bk, og, ch, smem, wait_upwind);
      // Every thread's coherent face stores must have completed before the flag.
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (threadIdx.x == 0)
        publish_epoch(flags + (bk * a.nty + bj) * a.ntx + bi, a.epoch);
    }
  }
}

// Batched persistent launch: the tasks of many 3D problems (same kernel shape
// and quadrature, any extents) in one ticket order, built on the host so that
// every task's upwind tasks hold smaller tickets (per problem, diagonal order;
// across problems, interleaved by diagonal index).  Per-problem arguments are
// read from a device table by the uniform problem index of the task.
template <typename R, int SCHEME, int SHAPE>
__global__ __launch_bounds__(kShapes[SHAPE].NT)
__attribute__((amdgpu_waves_per_eu(waves_per_simd<R, 3, SHAPE>()))) void
brick_batch_kernel(BatchArgs b)
{
  constexpr Shape S = kShapes[SHAPE];
  static_assert(S.dims == 3, "the batched sweep is 3D");
  using L = Layout<R, 3, S.B, S.DL, S.NDL, S.KO>;
  extern __shared__ __align__(16) unsigned char smem[];
  int * s_task = reinterpret_cast<int *>(smem + L::TASK_OFF);
  // Every problem of a batch shares the brick edge and so the plane table.
  load_plane_table<R, 3, S.B, S.DL, S.NDL, S.KO, S.NT>(b.probs[0], smem);
  // This work-group's ticket queue (see brick_flow_kernel): its tasks are
  // tasks[qoff[queue] ..) and their upwind tasks sit earlier in the same queue.
  const int queue = (int)(blockIdx.x % b.nqueues);
  unsigned long long * ticket = b.ticket + (size_t)queue * kTicketStride;
  const unsigned long long base = b.qbase[queue];
  const int first = b.qoff[queue];
  const unsigned long long count = (unsigned long long)(b.qoff[queue + 1] - first);
  while (true)
  {
    __syncthreads();  // see brick_flow_kernel
    if (threadIdx.x == 0)
    {
      const unsigned long long u = atomicAdd(ticket, 1ull) - base;
      *s_task = u < count ? first + (int)u : -1;
    }
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(*s_task);
    if (t < 0)
      break;
    const uint32_t code = b.tasks[t];
    const int p = (int)(code >> 20);
    const int br = (int)((code >> 6) & (kBatchMaxBricks - 1));
    const int r = (int)(code & (kBatchMaxPairs - 1));
    const BrickArgs a = b.probs[p];
    const int nbricks = a.ntx * a.nty * a.ntz;
    const int og = r / a.nchunks, ch = r % a.nchunks;
    const int bi = a.bricks[3 * br + 0], bj = a.bricks[3 * br + 1], bk = a.bricks[3 * br + 2];
    uint32_t * flags = a.flags + (size_t)r * nbricks;
    auto wait_upwind = [&]() {
      if (threadIdx.x == 0)
      {
        if (bi > 0)
          wait_epoch(flags + (bk * a.nty + bj) * a.ntx + (bi - 1), b.epoch, b.error,
                     b.wait_ticks);
        if (bj > 0)
          wait_epoch(flags + (bk * a.nty + (bj - 1)) * a.ntx + bi, b.epoch, b.error,
                     b.wait_ticks);
        if (bk > 0)
          wait_epoch(flags + ((bk - 1) * a.nty + bj) * a.ntx + bi, b.epoch, b.error,
                     b.wait_ticks);
      }
    };
    sweep_brick<R, 3, SCHEME, false, S.B, S.DL, S.NDL, S.KO, S.NT, true>(a, bi, bj, bk, og, ch,
                                                                         smem, wait_upwind);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
      publish_epoch(flags + (bk * a.nty + bj) * a.ntx + bi, b.epoch);
  }
}

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

template <typename R, int DIM, int SHAPE>
constexpr size_t
shmem_bytes()
{
  constexpr Shape S = kShapes[SHAPE];
  constexpr size_t bytes = Layout<R, DIM, S.B, S.DL, S.NDL, S.KO>::BYTES;
  static_assert(bytes <= 160 * 1024, "a kernel shape must fit the 160 KiB LDS of a CDNA4 CU");
  return bytes;
}

// Launch (op 0) or occupancy query (op 1) of one instantiated kernel.
template <typename R, int DIM, int SCHEME, bool EXACT, int SHAPE>
int
run_shape(const BrickArgs & a, int op, int schedule, int grid, hipStream_t s, int * per_cu)
{
  constexpr Shape S = kShapes[SHAPE];
  constexpr size_t shmem = shmem_bytes<R, DIM, SHAPE>();
  if (schedule == 0)
  {
    auto k = brick_sweep_kernel<R, DIM, SCHEME, EXACT, SHAPE>;
    static bool ready = false;
    int rc = prepare(k, shmem, ready);
    if (rc)
      return rc;
    if (op == 1)
    {
      hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, S.NT, shmem);
      return e == hipSuccess ? DOMA_OK : hip_fail(e, "occupancy query");
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(S.NT), shmem, s, a);
  }
  else if (schedule == 1)
  {
    auto k = brick_flow_kernel<R, DIM, SCHEME, EXACT, SHAPE>;
    static bool ready = false;
    int rc = prepare(k, shmem, ready);
    if (rc)
      return rc;
    if (op == 1)
    {
      hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, S.NT, shmem);
      return e == hipSuccess ? DOMA_OK : hip_fail(e, "occupancy query");
    }
    if (a.ev_start || a.ev_stop)
      hipExtLaunchKernelGGL(k, dim3(grid), dim3(S.NT), shmem, s, a.ev_start, a.ev_stop, 0, a);
    else
      hipLaunchKernelGGL(k, dim3(grid), dim3(S.NT), shmem, s, a);
  }
  else if constexpr (DIM == 3)
  {
    auto k = brick_pencil_kernel<R, DIM, SCHEME, EXACT, SHAPE>;
    static bool ready = false;
    int rc = prepare(k, shmem, ready);
    if (rc)
      return rc;
    if (op == 1)
    {
      hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, S.NT, shmem);
      return e == hipSuccess ? DOMA_OK : hip_fail(e, "occupancy query");
    }
    if (a.ev_start || a.ev_stop)
      hipExtLaunchKernelGGL(k, dim3(grid), dim3(S.NT), shmem, s, a.ev_start, a.ev_stop, 0, a);
    else
      hipLaunchKernelGGL(k, dim3(grid), dim3(S.NT), shmem, s, a);
  }
  else
    return fail(DOMA_ERR_INVALID, "the pencil schedule needs a 3D kernel shape");
  return DOMA_OK;
}

// Per-(precision, scheme) dispatch over the shape table; defined once per
// instantiation unit.
template <typename R, int SCHEME>
int dispatch_shapes(const BrickArgs & a, int exact, int shape, int op, int schedule, int grid,
                    hipStream_t s, int * per_cu);

template <typename R, int SCHEME, bool EXACT, int... IDS>
int
dispatch_seq(const BrickArgs & a, int shape, int op, int schedule, int grid, hipStream_t s,
             int * per_cu, std::integer_sequence<int, IDS...>)
{
  int rc = DOMA_ERR_INVALID;
  bool hit = false;
  (void)((shape == IDS
              ? (hit = true,
                 rc = run_shape<R, kShapes[IDS].dims, SCHEME, EXACT, IDS>(a, op, schedule, grid, s,
                                                                          per_cu),
                 true)
              : false) ||
         ...);
  if (!hit)
    return fail(DOMA_ERR_INVALID, "unknown sweep kernel shape");
  return rc;
}
// Batched launch (op 0) or occupancy query (op 1) for one shape; 3D shapes only.
template <typename R, int SCHEME, int SHAPE>
int
run_batch_shape(const BatchArgs & b, int op, int grid, hipStream_t s, int * per_cu)
{
  if constexpr (kShapes[SHAPE].dims != 3)
  {
    return fail(DOMA_ERR_INVALID, "the batched sweep needs a 3D kernel shape");
  }
  else
  {
    constexpr Shape S = kShapes[SHAPE];
    constexpr size_t shmem = shmem_bytes<R, 3, SHAPE>();
    auto k = brick_batch_kernel<R, SCHEME, SHAPE>;
    static bool ready = false;
    int rc = prepare(k, shmem, ready);
    if (rc)
      return rc;
    if (op == 1)
    {
      hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, S.NT, shmem);
      return e == hipSuccess ? DOMA_OK : hip_fail(e, "occupancy query");
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(S.NT), shmem, s, b);
    return DOMA_OK;
  }
}

template <typename R, int SCHEME>
int dispatch_batch_shapes(const BatchArgs & b, int shape, int op, int grid, hipStream_t s,
                          int * per_cu);

template <typename R, int SCHEME, int... IDS>
int
dispatch_batch_seq(const BatchArgs & b, int shape, int op, int grid, hipStream_t s, int * per_cu,
                   std::integer_sequence<int, IDS...>)
{
  int rc = DOMA_ERR_INVALID;
  bool hit = false;
  (void)((shape == IDS ? (hit = true, rc = run_batch_shape<R, SCHEME, IDS>(b, op, grid, s, per_cu),
                          true)
                       : false) ||
         ...);
  if (!hit)
    return fail(DOMA_ERR_INVALID, "unknown sweep kernel shape");
  return rc;
}
} // namespace sweepk
} // namespace doma

// Defines dispatch_shapes<R, SCHEME> for one (precision, scheme) pair.
#define DOMA_SWEEP_INSTANTIATE(R, SCHEME)                                                          \
  namespace doma                                                                                   \
  {                                                                                                \
  namespace sweepk                                                                                 \
  {                                                                                                \
  template <>                                                                                      \
  int dispatch_shapes<R, SCHEME>(const BrickArgs & a, int exact, int shape, int op, int schedule, \
                                 int grid, hipStream_t s, int * per_cu)                            \
  {                                                                                                \
    using seq = std::make_integer_sequence<int, kNumShapes>;                                       \
    return exact ? dispatch_seq<R, SCHEME, true>(a, shape, op, schedule, grid, s, per_cu, seq{})   \
                 : dispatch_seq<R, SCHEME, false>(a, shape, op, schedule, grid, s, per_cu, seq{}); \
  }                                                                                                \
  template <>                                                                                      \
  int dispatch_batch_shapes<R, SCHEME>(const BatchArgs & b, int shape, int op, int grid,           \
                                       hipStream_t s, int * per_cu)                                \
  {                                                                                                \
    using seq = std::make_integer_sequence<int, kNumShapes>;                                       \
    return dispatch_batch_seq<R, SCHEME>(b, shape, op, grid, s, per_cu, seq{});                    \
  }                                                                                                \
  }                                                                                                \
  }
