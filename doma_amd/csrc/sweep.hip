// sweep.hip -- brick-wavefront transport sweep for CDNA4 (gfx950).
//
// Decomposition (2D and 3D):
//   * The grid is cut into B^3 bricks (B^2 tiles in 2D).  For every octant the
//     bricks are visited in diagonals bi+bj+bk = s of the octant's own sweep
//     orientation; one launch processes every brick of diagonal s for all
//     octants and direction sets at once (vacuum boundaries make the octants
//     independent, TransportSolver3D.cpp:565-607 / CartesianCell3D.cpp:31-38).
//   * One work-group owns one (brick, octant, direction set).  Inside the brick
//     the cells are processed by hyperplanes li+lj+lk = p; the cells of a plane
//     are independent and run concurrently.  Threads are (cell slot, direction):
//     DL consecutive lanes carry the directions of one cell, so the per-cell
//     source and cross section are wave-uniform per lane group.
//   * The upwind face fluxes of the brick stay in LDS for the whole brick
//     (B^2 x-, y- and z-faces per direction, updated in place); only the brick's
//     inflow/outflow faces cross HBM, through per-direction-set face planes.
//   * The scalar flux of a cell is reduced over the DL direction lanes with a
//     fixed xor butterfly and stored once per (cell, direction set); the
//     direction sets are summed in a fixed order afterwards, so results are
//     reproducible run to run (no floating-point atomics).
//
// Cell solves: TWDiamondDifference3D.h:15-92, DiamondDifference3D.h:14-49,
// TWDiamondDifference2D.h:15-76, DiamondDifference2D.h:14-45,
// DiamondDifference1D.h:14-46.  The "exact" variants keep the reference's
// expression order with IEEE divides and no contraction; the default variants
// hoist the weight reciprocals (1/a = d/(d-n) with the same clamp) and use FMA.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace doma
{
namespace
{
// ------------------------------------------------------------------------- //
// Cell solves                                                                //
// ------------------------------------------------------------------------- //

template <typename R>
__device__ __forceinline__ R
twdd3d_exact(R q, R sig, R lx, R ly, R lz, R amu, R aeta, R axi, R & x, R & y, R & z)
{
#pragma clang fp contract(off)
  const R theta = R(0.9), theta_s = R(0.9);
  const R A = ly * lz;
  const R B = lx * lz;
  const R C = lx * ly;
  const R V = lx * ly * lz;
  R a = R(0.5);
  if (x > R(1e-8))
  {
    a = q * V * theta_s + (aeta * B * y + axi * C * z) * theta + amu * A * x;
    a /= (sig * V + R(2.0) * aeta * B + R(2.0) * axi * C) * x;
    a = R(-1.0) * (a - R(1.0));
  }
  R b = R(0.5);
  if (y > R(1e-8))
  {
    b = q * V * theta_s + (amu * A * x + axi * C * z) * theta + aeta * B * y;
    b /= (sig * V + R(2.0) * amu * A + R(2.0) * axi * C) * y;
    b = R(-1.0) * (b - R(1.0));
  }
  R c = R(0.5);
  if (z > R(1e-8))
  {
    c = q * V * theta_s + (amu * A * x + aeta * B * y) * theta + axi * C * z;
    c /= (sig * V + R(2.0) * amu * A + R(2.0) * aeta * B) * z;
    c = R(-1.0) * (c - R(1.0));
  }
  a = a < R(0.5) ? R(0.5) : a;
  a = a > R(1.0) ? R(1.0) : a;
  b = b < R(0.5) ? R(0.5) : b;
  b = b > R(1.0) ? R(1.0) : b;
  c = c < R(0.5) ? R(0.5) : c;
  c = c > R(1.0) ? R(1.0) : c;
  R psi = q * V + (amu * A / a * x) + (aeta * B / b * y) + (axi * C / c * z);
  psi /= sig * V + (amu * A / a) + (aeta * B / b) + (axi * C / c);
  x = (psi / a) - (((R(1.0) - a) / a) * x);
  y = (psi / b) - (((R(1.0) - b) / b) * y);
  z = (psi / c) - (((R(1.0) - c) / c) * z);
  return psi;
}

// Inverse TWDD weight 1/a for a = 1 - n/d clamped to [0.5, 1]: a < 0.5 <=> n > d/2
// and a > 1 <=> n < 0 (d > 0 whenever the upwind flux exceeds 1e-8).
template <typename R>
__device__ __forceinline__ R
inv_weight(R n, R d)
{
  R r = d / (d - n);
  r = n < R(0) ? R(1) : r;
  r = n >= R(0.5) * d ? R(2) : r;
  return r;
}

template <typename R>
__device__ __forceinline__ R
twdd3d_fast(R qV, R sV, R mA, R eB, R xC, R & x, R & y, R & z)
{
  const R th = R(0.9);
  const R qVs = qV * th;
  const R mAx = mA * x, eBy = eB * y, xCz = xC * z;
  R ia = R(2), ib = R(2), ic = R(2);
  if (x > R(1e-8))
    ia = inv_weight(qVs + (eBy + xCz) * th + mAx, (sV + R(2) * (eB + xC)) * x);
  if (y > R(1e-8))
    ib = inv_weight(qVs + (mAx + xCz) * th + eBy, (sV + R(2) * (mA + xC)) * y);
  if (z > R(1e-8))
    ic = inv_weight(qVs + (mAx + eBy) * th + xCz, (sV + R(2) * (mA + eB)) * z);
  const R wa = mA * ia, wb = eB * ib, wc = xC * ic;
  const R psi = (qV + wa * x + wb * y + wc * z) / (sV + wa + wb + wc);
  x = ia * psi - (ia - R(1)) * x;
  y = ib * psi - (ib - R(1)) * y;
  z = ic * psi - (ic - R(1)) * z;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd3d_exact(R q, R sig, R lx, R ly, R lz, R amu, R aeta, R axi, R & x, R & y, R & z)
{
#pragma clang fp contract(off)
  R psi = q + (R(2.0) * x * amu / lx) + (R(2.0) * y * aeta / ly) + (R(2.0) * z * axi / lz);
  psi /= (R(2.0) * amu / lx) + (R(2.0) * aeta / ly) + (R(2.0) * axi / lz) + sig;
  x = R(2.0) * psi - x;
  y = R(2.0) * psi - y;
  z = R(2.0) * psi - z;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd3d_fast(R q, R sig, R cx, R cy, R cz, R & x, R & y, R & z)
{
  const R psi = (q + cx * x + cy * y + cz * z) / (cx + cy + cz + sig);
  x = R(2) * psi - x;
  y = R(2) * psi - y;
  z = R(2) * psi - z;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
twdd2d_exact(R q, R sig, R lx, R ly, R amu, R aeta, R axi, R & x, R & y)
{
#pragma clang fp contract(off)
  const R theta = R(0.9), theta_s = R(0.9);
  const R A = ly;
  const R B = lx;
  const R V = lx * ly;
  R a = R(0.5);
  if (x > R(1e-8))
  {
    a = q * V * theta_s + (aeta * B * y) * theta + amu * A * x;
    a /= (sig * V + R(2.0) * aeta * B) * x;
    a = R(-1.0) * (a - R(1.0));
  }
  R b = R(0.5);
  if (y > R(1e-8))
  {
    b = q * V * theta_s + (amu * A * x + axi) * theta + aeta * B * y;
    b /= (sig * V + R(2.0) * amu * A) * y;
    b = R(-1.0) * (b - R(1.0));
  }
  a = a < R(0.5) ? R(0.5) : a;
  a = a > R(1.0) ? R(1.0) : a;
  b = b < R(0.5) ? R(0.5) : b;
  b = b > R(1.0) ? R(1.0) : b;
  R psi = q * V + (amu * A / a * x) + (aeta * B / b * y);
  psi /= sig * V + (amu * A / a) + (aeta * B / b);
  x = (psi / a) - (((R(1.0) - a) / a) * x);
  y = (psi / b) - (((R(1.0) - b) / b) * y);
  return psi;
}

template <typename R>
__device__ __forceinline__ R
twdd2d_fast(R qV, R sV, R mA, R eB, R axi, R & x, R & y)
{
  const R th = R(0.9);
  const R qVs = qV * th;
  const R mAx = mA * x, eBy = eB * y;
  R ia = R(2), ib = R(2);
  if (x > R(1e-8))
    ia = inv_weight(qVs + eBy * th + mAx, (sV + R(2) * eB) * x);
  if (y > R(1e-8))
    ib = inv_weight(qVs + (mAx + axi) * th + eBy, (sV + R(2) * mA) * y);
  const R wa = mA * ia, wb = eB * ib;
  const R psi = (qV + wa * x + wb * y) / (sV + wa + wb);
  x = ia * psi - (ia - R(1)) * x;
  y = ib * psi - (ib - R(1)) * y;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd2d_exact(R q, R sig, R lx, R ly, R amu, R aeta, R & x, R & y)
{
#pragma clang fp contract(off)
  R psi = q + (R(2.0) * x * amu / lx) + (R(2.0) * y * aeta / ly);
  psi /= (R(2.0) * amu / lx) + (R(2.0) * aeta / ly) + sig;
  x = R(2.0) * psi - x;
  y = R(2.0) * psi - y;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd1d(R q, R sig, R lx, R amu, R & x)
{
#pragma clang fp contract(off)
  R psi = q + (R(2.0) * x * amu / lx);
  psi /= (R(2.0) * amu / lx) + sig;
  x = R(2.0) * psi - x;
  return psi;
}

// Fixed-order butterfly sum over groups of DL consecutive lanes; every lane of
// the group receives the same (bitwise) result.
template <int DL>
__device__ __forceinline__ double
group_sum(double v)
{
#pragma unroll
  for (int m = DL / 2; m >= 1; m >>= 1)
    v += __shfl_xor(v, m, DL);
  return v;
}

// ------------------------------------------------------------------------- //
// Brick sweep kernel                                                         //
// ------------------------------------------------------------------------- //

template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
__global__ __launch_bounds__(NT) void
brick_sweep_kernel(BrickArgs a)
{
  constexpr int CPP = NT / DL;                  // cells per pass
  constexpr int FS = (DIM == 3 ? B * B : B);    // face slots per axis
  constexpr int NPLANES = DIM == 3 ? 3 * B - 2 : 2 * B - 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R * sx = reinterpret_cast<R *>(smem_raw);
  R * sy = sx + FS * DL;
  R * sz = sy + FS * DL;  // 3D only

  const int ds = blockIdx.x % a.ndirsets;
  const int br = a.diag_begin + blockIdx.x / a.ndirsets;
  const int oct = ds / a.nchunks;
  const int chunk = ds - oct * a.nchunks;
  const int bi = a.bricks[3 * br + 0];
  const int bj = a.bricks[3 * br + 1];
  const int bk = a.bricks[3 * br + 2];
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
  // Global face plane offsets (oriented coordinates, padded extents).
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

  // Inflow faces: the upwind brick's outflow, or vacuum (zero) on the boundary.
  for (int e = tid; e < FS * DL; e += NT)
  {
    const int s = e / DL, l = e - s * DL;
    const int u = DIM == 3 ? s / B : s;  // first in-plane coordinate
    const int v = DIM == 3 ? s % B : 0;  // second in-plane coordinate
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
  __syncthreads();

  for (int p = 0; p < NPLANES; ++p)
  {
    const int pe = a.plane_off[p + 1];
    for (int pos = a.plane_off[p] + slot; pos < pe; pos += CPP)
    {
      const uint32_t pc = a.plane_cells[pos];
      const int li = pc & 0xff, lj = (pc >> 8) & 0xff, lk = (pc >> 16) & 0xff;
      const int oi = i0 + li, oj = j0 + lj, ok = k0 + lk;
      if (oi >= a.nx || oj >= a.ny || ok >= a.nz)
        continue;  // outside a partial brick: the whole lane group skips
      const int i = negx ? a.nx - 1 - oi : oi;
      const int j = negy ? a.ny - 1 - oj : oj;
      const int k = negz ? a.nz - 1 - ok : ok;
      const int c = (k * a.ny + j) * a.nx + i;
      const double qd = a.q[c];
      const double sd = a.sig[c];
      R psi;
      if (DIM == 3)
      {
        R & xf = sx[(lj * B + lk) * DL + dd];
        R & yf = sy[(li * B + lk) * DL + dd];
        R & zf = sz[(li * B + lj) * DL + dd];
        R x = xf, y = yf, z = zf;
        if (SCHEME == DOMA_SCHEME_TW_DD)
        {
          if (EXACT)
            psi = twdd3d_exact<R>(R(qd), R(sd), R(a.lx[i]), R(a.ly[j]), R(a.lz[k]), amu, aeta, axi,
                                  x, y, z);
          else
          {
            const double lxd = a.lx[i], lyd = a.ly[j], lzd = a.lz[k];
            const double V = lxd * lyd * lzd;
            psi = twdd3d_fast<R>(R(qd * V), R(sd * V), amu * R(lyd * lzd), aeta * R(lxd * lzd),
                                 axi * R(lxd * lyd), x, y, z);
          }
        }
        else
        {
          if (EXACT)
            psi = dd3d_exact<R>(R(qd), R(sd), R(a.lx[i]), R(a.ly[j]), R(a.lz[k]), amu, aeta, axi, x,
                                y, z);
          else
            psi = dd3d_fast<R>(R(qd), R(sd), R(2) * amu * R(a.rlx[i]), R(2) * aeta * R(a.rly[j]),
                               R(2) * axi * R(a.rlz[k]), x, y, z);
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
            psi = twdd2d_exact<R>(R(qd), R(sd), R(a.lx[i]), R(a.ly[j]), amu, aeta, axi, x, y);
          else
          {
            const double lxd = a.lx[i], lyd = a.ly[j];
            const double V = lxd * lyd;
            psi = twdd2d_fast<R>(R(qd * V), R(sd * V), amu * R(lyd), aeta * R(lxd), axi, x, y);
          }
        }
        else
        {
          if (EXACT)
            psi = dd2d_exact<R>(R(qd), R(sd), R(a.lx[i]), R(a.ly[j]), amu, aeta, x, y);
          else
          {
            R z = R(0);
            psi = dd3d_fast<R>(R(qd), R(sd), R(2) * amu * R(a.rlx[i]), R(2) * aeta * R(a.rly[j]),
                               R(0), x, y, z);
          }
        }
        xf = x;
        yf = y;
      }
      double s = valid ? w * static_cast<double>(psi) : 0.0;
      s = group_sum<DL>(s);
      if (dd == 0)
        a.part[(size_t)ds * a.N + c] = s;
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

template <typename R, int DIM, int SCHEME, bool EXACT, int B, int DL, int NT>
int
launch_one(const BrickArgs & a, hipStream_t s)
{
  constexpr int FS = (DIM == 3 ? B * B : B);
  const size_t shmem = sizeof(R) * FS * DL * (DIM == 3 ? 3 : 2);
  auto k = brick_sweep_kernel<R, DIM, SCHEME, EXACT, B, DL, NT>;
  static bool attr_set = false;
  if (!attr_set)
  {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess)
      return hip_fail(e, "hipFuncSetAttribute(brick_sweep_kernel)");
    attr_set = true;
  }
  const unsigned grid = (unsigned)a.diag_count * (unsigned)a.ndirsets;
  hipLaunchKernelGGL(k, dim3(grid), dim3(NT), shmem, s, a);
  return DOMA_OK;
}

template <typename R, int DIM, int SCHEME, bool EXACT, int B, int NT>
int
dispatch_dl(const BrickArgs & a, int dl, hipStream_t s)
{
  switch (dl)
  {
    case 8: return launch_one<R, DIM, SCHEME, EXACT, B, 8, NT>(a, s);
    case 16: return launch_one<R, DIM, SCHEME, EXACT, B, 16, NT>(a, s);
    case 32: return launch_one<R, DIM, SCHEME, EXACT, B, 32, NT>(a, s);
    case 64: return launch_one<R, DIM, SCHEME, EXACT, B, 64, NT>(a, s);
  }
  return fail(DOMA_ERR_INVALID, "unsupported direction-lane width");
}

template <typename R, int DIM, int SCHEME, bool EXACT>
int
dispatch_shape(const BrickArgs & a, int brick, int dl, int threads, hipStream_t s)
{
  if (DIM == 3)
  {
    if (brick == 4 && threads == 256) return dispatch_dl<R, DIM, SCHEME, EXACT, 4, 256>(a, dl, s);
    if (brick == 8 && threads == 256) return dispatch_dl<R, DIM, SCHEME, EXACT, 8, 256>(a, dl, s);
    if (brick == 8 && threads == 512) return dispatch_dl<R, DIM, SCHEME, EXACT, 8, 512>(a, dl, s);
    if (brick == 4 && threads == 128) return dispatch_dl<R, DIM, SCHEME, EXACT, 4, 128>(a, dl, s);
  }
  else
  {
    if (brick == 8 && threads == 256) return dispatch_dl<R, DIM, SCHEME, EXACT, 8, 256>(a, dl, s);
    if (brick == 16 && threads == 256) return dispatch_dl<R, DIM, SCHEME, EXACT, 16, 256>(a, dl, s);
  }
  return fail(DOMA_ERR_INVALID, "unsupported brick / work-group shape");
}

template <typename R, int DIM>
int
dispatch_scheme(const BrickArgs & a, int scheme, int exact, int brick, int dl, int threads,
                hipStream_t s)
{
  if (scheme == DOMA_SCHEME_TW_DD)
    return exact ? dispatch_shape<R, DIM, DOMA_SCHEME_TW_DD, true>(a, brick, dl, threads, s)
                 : dispatch_shape<R, DIM, DOMA_SCHEME_TW_DD, false>(a, brick, dl, threads, s);
  return exact ? dispatch_shape<R, DIM, DOMA_SCHEME_DD, true>(a, brick, dl, threads, s)
               : dispatch_shape<R, DIM, DOMA_SCHEME_DD, false>(a, brick, dl, threads, s);
}
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
                   int dl, int threads, hipStream_t s)
{
  if (!brick_variant_exists(dims, brick, dl, threads))
    return fail(DOMA_ERR_INVALID, "no sweep kernel for this brick / lane / work-group shape");
  if (dims == 3)
    return precision == DOMA_FP32_SWEEP
               ? dispatch_scheme<float, 3>(a, scheme, exact, brick, dl, threads, s)
               : dispatch_scheme<double, 3>(a, scheme, exact, brick, dl, threads, s);
  return precision == DOMA_FP32_SWEEP
             ? dispatch_scheme<float, 2>(a, scheme, exact, brick, dl, threads, s)
             : dispatch_scheme<double, 2>(a, scheme, exact, brick, dl, threads, s);
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
