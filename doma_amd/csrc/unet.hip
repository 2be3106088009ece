// unet.hip -- inference kernels of the learned SI correction (3-level 3D U-Net,
// BF16 activations and weights, FP32 accumulation on MFMA).
//
// Convolutions are implicit GEMMs: rows = output voxels (16 consecutive voxels
// per MFMA tile), columns = output channels (16 per tile), K = tap * cin_p + ci
// over the 27 taps of a 3x3x3 stencil.  Each lane of a wave feeds the
// v_mfma_f32_16x16x32_bf16 A operand with 8 consecutive channels of one tap of
// its voxel (one 16-byte load, channels-last layout) and the B operand with 8
// consecutive K entries of its output channel (weights packed K-contiguous).
// The decoder's concat(skip, upsample(coarse)) input is gathered on the fly: K
// slots below cin_a read the skip tensor at the voxel, the rest read the coarse
// tensor at the parent voxel, so neither the upsampled nor the concatenated
// tensor is ever stored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>

#include "internal.h"
#include "unet.h"

namespace doma
{
namespace net
{
namespace
{
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

constexpr int RT = 256;

__host__ __device__ inline uint16_t
f2bf(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ inline float
bf2f(uint16_t h)
{
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__device__ inline v8bf
load8(const uint16_t * p)
{
  const v4u u = *reinterpret_cast<const v4u *>(p);
  return __builtin_bit_cast(v8bf, u);
}

struct ConvArgs
{
  const uint16_t * a;  // full-resolution input [V][ca]
  int ca;              // its channel stride (= cin_a_p)
  const uint16_t * b;  // coarse input [V/8][cb], read at the parent voxel (or null)
  int cb;
  int D, H, W, V;      // D counts every sample of a batch stacked along z
  int Ds;              // depth of one sample (zero padding at its z faces)
  const uint16_t * w;  // [cout16][kpad]
  const float * bias;  // [cout]
  int cout, kpad, cin_p;
  uint16_t * out;      // [V][out_stride]
  int out_stride;
  int relu;
  int four;            // only input channels 0..3 are non-zero (cin_p = 8, no coarse input)
};

template <int NT, int MT>
__global__ __launch_bounds__(RT) void
conv3_kernel(ConvArgs p)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const int base = (blockIdx.x * (RT / 64) + wave) * 16 * MT;
  const int H2 = p.H >> 1, W2 = p.W >> 1;
  int vz[MT], vzl[MT], vy[MT], vx[MT];
  bool vin[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
  {
    const int v = base + mt * 16 + row;
    vin[mt] = v < p.V;
    const int vv = vin[mt] ? v : 0;
    vx[mt] = vv % p.W;
    vy[mt] = (vv / p.W) % p.H;
    vz[mt] = vv / (p.W * p.H);
    vzl[mt] = vz[mt] % p.Ds;
  }
  v4f acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      acc[mt][nt] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < p.kpad; k0 += 32)
  {
    const int k = k0 + kg * 8;
    const int tap = k / p.cin_p;
    const int ci = k - tap * p.cin_p;
    v8bf bw[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      bw[nt] = load8(p.w + (size_t)(nt * 16 + row) * p.kpad + k);
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
      v4u raw = {0u, 0u, 0u, 0u};
      const int z = vz[mt] + dz, zl = vzl[mt] + dz, y = vy[mt] + dy, x = vx[mt] + dx;
      if (tap < 27 && vin[mt] && zl >= 0 && zl < p.Ds && y >= 0 && y < p.H && x >= 0 && x < p.W)
      {
        if (ci < p.ca)
          raw = *reinterpret_cast<const v4u *>(p.a + ((size_t)(z * p.H + y) * p.W + x) * p.ca + ci);
        else
          raw = *reinterpret_cast<const v4u *>(
              p.b + ((size_t)((z >> 1) * H2 + (y >> 1)) * W2 + (x >> 1)) * p.cb + (ci - p.ca));
      }
      const v8bf av = __builtin_bit_cast(v8bf, raw);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  // D layout: lane holds rows 4 * (lane / 16) + j of column lane % 16.
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
    {
      const int co = nt * 16 + row;
      if (co >= p.cout)
        continue;
      const float bv = p.bias[co];
#pragma unroll
      for (int j = 0; j < 4; ++j)
      {
        const int v = base + mt * 16 + kg * 4 + j;
        if (v >= p.V)
          continue;
        float f = acc[mt][nt][j] + bv;
        if (p.relu)
          f = fmaxf(f, 0.f);
        p.out[(size_t)v * p.out_stride + co] = f2bf(f);
      }
    }
}

// The same convolution with the input staged in LDS, CG = 8 channels at a time:
// a work-group computes an output tile of TZ x 8 x 16 voxels (rows of 16
// consecutive x voxels, TZ*8/4 rows per wave).  For each group of 8 input
// channels it copies the tile's input halo, (TZ+2) x 10 x 18 voxels, and the
// group's weight fragments into LDS (the decoder's coarse channels gathered from
// the parent voxel on the way), so each input value crosses the cache hierarchy
// once instead of once per tap, and then runs 7 k-steps of 4 taps x 8 channels
// with both operands read from LDS.  The next group's halo and weights are
// loaded into registers while the current group is multiplied.  The MFMA
// computes the transposed tile, out^T = W^T x in^T: the A operand is one weight
// fragment per k-step (16 output channels x 4 taps x 8 channels) and the B
// operand is the voxel row's 8 channels at one tap, so each lane ends up holding
// 4 consecutive output channels of one voxel and stores them as one 8-byte
// write.  A layer with 4 input channels (CG = 4) packs 8 taps x 4 channels per
// k-step instead (4 k-steps, no zero half): each lane's K slice is two taps of
// its voxel.  Used when a sample's depth Ds is a multiple of TZ (tiles never
// straddle samples) and the output channels come in fours.
constexpr int kTileX = 16, kTileY = 8;

template <int NT, int TZ, int CG>
__global__ __launch_bounds__(RT) void
conv3_tile_kernel(ConvArgs p)
{
  static_assert(CG == 8 || CG == 4, "channels per LDS group");
  using SV = std::conditional_t<CG == 8, v4u, v2u>;  // one voxel's group of channels
  constexpr int HX = kTileX + 2, HY = kTileY + 2, HZ = TZ + 2, HALO = HX * HY * HZ;
  constexpr int MT = TZ * kTileY / (RT / 64);  // row tiles per wave
  constexpr int SI = (HALO + RT - 1) / RT;     // halo voxels staged per thread
  constexpr int NTAP = CG == 8 ? 28 : 32;      // tap slots (27 real, the rest zero weight)
  constexpr int TPS = 32 / CG;                 // taps per k-step
  constexpr int WS = NTAP * 16 * NT;           // weight fragments (CG channels) per group
  constexpr int WI = (WS + RT - 1) / RT;       // weight fragments staged per thread
  // LDS: the halo [HALO][CG]; the weights, CG = 8: [28 taps][16 NT rows][8],
  // CG = 4: [16 NT rows][32 taps][4] (a lane's two taps are adjacent).
  __shared__ __align__(16) uint16_t tile[HALO * CG];
  __shared__ __align__(16) uint16_t wtile[WS * CG];
  __shared__ int tapoff[32];  // halo offset of each tap (zero-weight slots: any voxel)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const int ntx = (p.W + kTileX - 1) / kTileX, nty = (p.H + kTileY - 1) / kTileY;
  const int bx = (int)(blockIdx.x % ntx), by = (int)((blockIdx.x / ntx) % nty);
  const int bz = (int)(blockIdx.x / ((unsigned)ntx * nty));
  const int x0 = bx * kTileX, y0 = by * kTileY, z0 = bz * TZ;
  const int zs = z0 - z0 % p.Ds;  // first z of this tile's sample
  const int H2 = p.H >> 1, W2 = p.W >> 1;

  // Source voxels of this thread's halo slots (-1: zero padding or no slot).
  int va[SI], vb[SI];
#pragma unroll
  for (int i = 0; i < SI; ++i)
  {
    const int h = threadIdx.x + i * RT;
    const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
    const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
    const bool in = h < HALO && x >= 0 && x < p.W && y >= 0 && y < p.H && z >= zs &&
                    z < zs + p.Ds && z < p.D;
    va[i] = in ? (z * p.H + y) * p.W + x : -1;
    vb[i] = in ? ((z >> 1) * H2 + (y >> 1)) * W2 + (x >> 1) : -1;
  }
  // A channel group's halo and weight fragments, loaded into registers.
  SV stage[SI], wstage[WI];
  auto fetch = [&](int c0) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < SI; ++i)
    {
      SV raw = {};
      if (va[i] >= 0)
      {
        if (c0 < p.ca)
          raw = *reinterpret_cast<const SV *>(p.a + (size_t)va[i] * p.ca + c0);
        else
          raw = *reinterpret_cast<const SV *>(p.b + (size_t)vb[i] * p.cb + (c0 - p.ca));
      }
      stage[i] = raw;
    }
#pragma unroll
    for (int i = 0; i < WI; ++i)
    {
      const int e = threadIdx.x + i * RT;
      int tap, co;
      if constexpr (CG == 8)
      {
        tap = e / (16 * NT);
        co = e - tap * (16 * NT);
      }
      else
      {
        co = e / NTAP;
        tap = e - co * NTAP;
      }
      SV wr = {};
      if (e < WS && tap < 27)
        wr = *reinterpret_cast<const SV *>(p.w + (size_t)co * p.kpad + tap * p.cin_p + c0);
      wstage[i] = wr;
    }
  };

  if (threadIdx.x < 32)
  {
    const int t = threadIdx.x < 27 ? threadIdx.x : 13;
    tapoff[threadIdx.x] = ((t / 9) * HY + (t / 3) % 3) * HX + t % 3;
  }
  v4f acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      acc[mt][nt] = v4f{0.f, 0.f, 0.f, 0.f};

  const int cend = CG == 8 ? p.cin_p : 1;  // CG = 4: the one group of channels 0..3
  fetch(0);
  for (int c0 = 0; c0 < cend; c0 += 8)
  {
#pragma unroll
    for (int i = 0; i < SI; ++i)
      if (threadIdx.x + i * RT < HALO)
        *reinterpret_cast<SV *>(tile + (size_t)(threadIdx.x + i * RT) * CG) = stage[i];
#pragma unroll
    for (int i = 0; i < WI; ++i)
      if (threadIdx.x + i * RT < WS)
        *reinterpret_cast<SV *>(wtile + (size_t)(threadIdx.x + i * RT) * CG) = wstage[i];
    __syncthreads();
    // The next group's global loads stay in flight over this group's MFMAs.
    if (c0 + 8 < cend)
      fetch(c0 + 8);
    // NTAP / TPS k-steps, operands from LDS only.
#pragma unroll 1
    for (int t0 = 0; t0 < NTAP; t0 += TPS)
    {
      v8bf bw[NT];
      int hoff0, hoff1 = 0;
      if constexpr (CG == 8)
      {
        const int tap = t0 + kg;
        hoff0 = tapoff[tap];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          bw[nt] = load8(wtile + (size_t)(tap * 16 * NT + nt * 16 + row) * 8);
      }
      else
      {
        const int tap = t0 + 2 * kg;
        hoff0 = tapoff[tap];
        hoff1 = tapoff[tap + 1];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          bw[nt] = load8(wtile + (size_t)((nt * 16 + row) * NTAP + tap) * 4);
      }
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
      {
        const int r = wave * MT + mt, tz = r / kTileY, ty = r - tz * kTileY;
        const int hb = (tz * HY + ty) * HX + row;
        v8bf av;
        if constexpr (CG == 8)
          av = load8(tile + (size_t)(hb + hoff0) * 8);
        else
        {
          const v2u lo = *reinterpret_cast<const v2u *>(tile + (size_t)(hb + hoff0) * 4);
          const v2u hi = *reinterpret_cast<const v2u *>(tile + (size_t)(hb + hoff1) * 4);
          av = __builtin_bit_cast(v8bf, v4u{lo[0], lo[1], hi[0], hi[1]});
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[nt], av, acc[mt][nt], 0, 0, 0);
      }
    }
    __syncthreads();  // the next group's staging overwrites the halo
  }
  // D layout: lane holds output channels nt*16 + 4*kg + j (j = 0..3) of voxel
  // x0 + row of its row tile.
  const int x = x0 + row;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
  {
    const int co = nt * 16 + 4 * kg;
    if (co >= p.cout)
      continue;
    const v4f bv = {p.bias[co], p.bias[co + 1], p.bias[co + 2], p.bias[co + 3]};
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
    {
      const int r = wave * MT + mt, tz = r / kTileY, ty = r - tz * kTileY;
      const int z = z0 + tz, y = y0 + ty;
      if (z >= p.D || y >= p.H || x >= p.W)
        continue;
      v4f f = acc[mt][nt] + bv;
      if (p.relu)
        f = __builtin_elementwise_max(f, v4f{0.f, 0.f, 0.f, 0.f});
      const v2bf lo = __builtin_convertvector(v2f{f[0], f[1]}, v2bf);
      const v2bf hi = __builtin_convertvector(v2f{f[2], f[3]}, v2bf);
      v2u packed = {__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)};
      *reinterpret_cast<v2u *>(p.out + ((size_t)(z * p.H + y) * p.W + x) * p.out_stride + co) =
          packed;
    }
  }
}

// 2x2x2 average pooling, channels-last (C a multiple of 8): one thread per
// output voxel and group of 8 channels, 16-byte loads and store.
__global__ void
pool_kernel(const uint16_t * in, uint16_t * out, int D, int H, int W, int C)
{
  const int D2 = D >> 1, H2 = H >> 1, W2 = W >> 1, C8 = C >> 3;
  const size_t n = (size_t)D2 * H2 * W2 * C8;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int c = (int)(i % C8) * 8;
  const size_t v = i / C8;
  const int x = (int)(v % W2), y = (int)((v / W2) % H2), z = (int)(v / ((size_t)W2 * H2));
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int d = 0; d < 8; ++d)
  {
    const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
    const v4u u = *reinterpret_cast<const v4u *>(in + ((size_t)(zz * H + yy) * W + xx) * C + c);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      s[j] += bf2f((uint16_t)(u[j >> 1] >> (16 * (j & 1))));
  }
  v4u o;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    o[j] = (uint32_t)f2bf(0.125f * s[2 * j]) | ((uint32_t)f2bf(0.125f * s[2 * j + 1]) << 16);
  *reinterpret_cast<v4u *>(out + v * C + c) = o;
}

// Fixed-order partial sums of d^2 (RT threads, grid-stride).
__global__ __launch_bounds__(RT) void
sumsq_kernel(const double * d, int N, double * red)
{
  __shared__ double sh[RT];
  double s = 0.0;
  for (int c = blockIdx.x * RT + threadIdx.x; c < N; c += gridDim.x * RT)
    s += d[c] * d[c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = RT / 2; k > 0; k >>= 1)
  {
    if ((int)threadIdx.x < k)
      sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    red[blockIdx.x] = sh[0];
}

// scale[0] = rms of the first sum (sum d_k^2 over N cells), scale[1] = the
// extrapolation gain rho / (1 - rho), rho = rms(d_k) / rms(d_(k-1)) clamped to
// [0, 0.999] (red holds the two sets of block partial sums).
__global__ __launch_bounds__(RT) void
rms_final_kernel(const double * red, int nblocks, int N, double * scale)
{
  __shared__ double sh[2][RT];
  double s0 = 0.0, s1 = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += RT)
  {
    s0 += red[b];
    s1 += red[nblocks + b];
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  for (int k = RT / 2; k > 0; k >>= 1)
  {
    if ((int)threadIdx.x < k)
    {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + k];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
  {
    const double r = sqrt(sh[0][0] / N), rp = sqrt(sh[1][0] / N);
    scale[0] = r > 0.0 ? r : 1.0;
    double rho = rp > 0.0 ? r / rp : 0.0;
    rho = rho < 0.0 ? 0.0 : (rho > 0.999 ? 0.999 : rho);
    scale[1] = rho / (1.0 - rho);
  }
}

// Network input of group g: the last two increments scaled by the rms of the
// latest, log sigma_t and log 1/(1 - c_gg + 1e-4) with c_gg = sigma_s,gg / sigma_t.
__global__ void
features_kernel(const double * d, const double * dp, const double * scale, const int32_t * mat,
                const double * sigt, const double * sigs, int G, int g, int N, uint16_t * feat)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N)
    return;
  const int m = mat[c];
  const double st = sigt[(size_t)m * G + g];
  const double sgg = sigs[((size_t)m * G + g) * G + g];
  const double cgg = st > 0.0 ? sgg / st : 0.0;
  uint16_t * o = feat + (size_t)c * 8;
  o[0] = f2bf((float)(d[c] / scale[0]));
  o[1] = f2bf((float)(dp[c] / scale[0]));
  o[2] = f2bf((float)log(st > 1e-6 ? st : 1e-6));
  o[3] = f2bf((float)log(1.0 / (1.0 - cgg + 1e-4)));
  for (int k = CIN; k < 8; ++k)
    o[k] = 0;
}

__global__ void
features_f32_kernel(const float * in, int N, uint16_t * feat)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N)
    return;
  for (int k = 0; k < 8; ++k)
    feat[(size_t)c * 8 + k] = k < CIN ? f2bf(in[(size_t)c * CIN + k]) : 0;
}

// 1x1x1 head: y = w . h + b.  Raw output for the stand-alone forward pass.
// The voxel's channels (stride a multiple of 8) are read 8 at a time with
// 16-byte loads; the sum runs over channels 0..C-1 in order.
__device__ inline float
head_value(const uint16_t * h, int stride, int C, const float * wb, int c)
{
  float s = wb[C];
  const uint16_t * row = h + (size_t)c * stride;
  for (int k0 = 0; k0 < C; k0 += 8)
  {
    const v4u u = *reinterpret_cast<const v4u *>(row + k0);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (k0 + j < C)
        s += wb[k0 + j] * bf2f((uint16_t)(u[j >> 1] >> (16 * (j & 1))));
  }
  return s;
}

__global__ void
head_raw_kernel(const uint16_t * h, int stride, int C, const float * wb, int N, float * out)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < N)
    out[c] = head_value(h, stride, C, wb, c);
}

// Correction: e = gain * (d_k + s * y) -- the geometric-series extrapolation
// gain * d_k of the remaining error plus the network's refinement of it (a zero
// network leaves the bare extrapolation).
__global__ void
head_corr_kernel(const uint16_t * h, int stride, int C, const float * wb, const double * scale,
                 const double * dlast, int N, double * out)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N)
    return;
  out[c] = scale[1] * (dlast[c] + scale[0] * (double)head_value(h, stride, C, wb, c));
}

inline unsigned
nblk(size_t n, int t)
{
  return (unsigned)((n + t - 1) / t);
}

template <int TZ>
unsigned
conv_tiles(const ConvArgs & p)
{
  return (unsigned)(((p.W + kTileX - 1) / kTileX) * ((p.H + kTileY - 1) / kTileY) *
                    ((p.D + TZ - 1) / TZ));
}

template <int NT, int TZ>
int
launch_tile(const ConvArgs & p, hipStream_t s)
{
  if (p.four)
    hipLaunchKernelGGL((conv3_tile_kernel<NT, TZ, 4>), dim3(conv_tiles<TZ>(p)), dim3(RT), 0, s, p);
  else
    hipLaunchKernelGGL((conv3_tile_kernel<NT, TZ, 8>), dim3(conv_tiles<TZ>(p)), dim3(RT), 0, s, p);
  return DOMA_OK;
}

template <int TZ>
int
launch_conv_tile(const ConvArgs & p, int cout16, hipStream_t s)
{
  if (cout16 == 16)
    return launch_tile<1, TZ>(p, s);
  if (cout16 == 32)
    return launch_tile<2, TZ>(p, s);
  if (cout16 == 64)
    return launch_tile<4, TZ>(p, s);
  return fail(DOMA_ERR_INVALID, "unsupported layer width");
}

int
launch_conv(const ConvArgs & p, int cout16, hipStream_t s)
{
  // The LDS-tiled kernel when the K layout is 27 taps of cin_p channels and the
  // output channels (and the output row stride) come in fours, with tiles of 4,
  // 2 or 1 z-planes: the deepest whose tiles do not straddle a sample and still
  // number at least 256 (one per CU); the direct kernel otherwise.
  const bool tiled = p.kpad >= 27 * p.cin_p && p.cout % 4 == 0 && p.out_stride % 4 == 0 &&
                     (!p.four || (p.cin_p == 8 && p.cb == 0));
  if (tiled)
  {
    if (p.Ds % 4 == 0 && conv_tiles<4>(p) >= 256)
      return launch_conv_tile<4>(p, cout16, s);
    if (p.Ds % 2 == 0 && conv_tiles<2>(p) >= 256)
      return launch_conv_tile<2>(p, cout16, s);
    return launch_conv_tile<1>(p, cout16, s);
  }
  constexpr int MT = 2;
  const unsigned grid = nblk(p.V, (RT / 64) * 16 * MT);
  if (cout16 == 16)
    hipLaunchKernelGGL((conv3_kernel<1, MT>), dim3(grid), dim3(RT), 0, s, p);
  else if (cout16 == 32)
    hipLaunchKernelGGL((conv3_kernel<2, MT>), dim3(grid), dim3(RT), 0, s, p);
  else if (cout16 == 64)
    hipLaunchKernelGGL((conv3_kernel<4, MT>), dim3(grid), dim3(RT), 0, s, p);
  else
    return fail(DOMA_ERR_INVALID, "unsupported layer width");
  return DOMA_OK;
}

int
conv(const PackedLayer & L, const uint16_t * a, const uint16_t * b, int D, int H, int W,
     uint16_t * out, int out_stride, hipStream_t s, int Ds = 0, bool four = false)
{
  ConvArgs p{};
  p.four = four;
  p.a = a;
  p.ca = L.cin_a_p;
  p.b = b;
  p.cb = L.cin_b_p;
  p.D = D;
  p.H = H;
  p.W = W;
  p.V = D * H * W;
  p.Ds = Ds > 0 ? Ds : D;
  p.w = L.w;
  p.bias = L.bias;
  p.cout = L.cout;
  p.kpad = L.kpad;
  p.cin_p = L.cin_a_p + L.cin_b_p;
  p.out = out;
  p.out_stride = out_stride;
  p.relu = L.relu;
  return launch_conv(p, L.cout16, s);
}

int
pool(const uint16_t * in, uint16_t * out, int D, int H, int W, int C, hipStream_t s)
{
  if (C % 8)
    return fail(DOMA_ERR_INVALID, "pooling needs a channel stride that is a multiple of 8");
  const size_t n = (size_t)(D / 2) * (H / 2) * (W / 2) * (C / 8);
  hipLaunchKernelGGL(pool_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, in, out, D, H, W, C);
  return DOMA_OK;
}

// Layers 0-9 on the features in n.feat; leaves the last hidden layer in n.d1b.
// Ds: depth of one sample when a batch is stacked along z (0: one sample).
int
body(Net & n, hipStream_t s, int Ds = 0)
{
  const int D = n.D, H = n.H, W = n.W;
  const int D2 = D / 2, H2 = H / 2, W2 = W / 2, D4 = D / 4, H4 = H / 4, W4 = W / 4;
  const int s1 = pad8(C1), s2 = pad8(C2), s3 = pad8(C3);
  const int z1 = Ds > 0 ? Ds : D, z2 = z1 / 2, z4 = z1 / 4;
  int rc;
  static_assert(CIN == 4, "layer 0 reads the 4 feature channels of an 8-channel stride");
  if ((rc = conv(n.layer[0], n.feat, nullptr, D, H, W, n.e1a, s1, s, z1, true)) ||
      (rc = conv(n.layer[1], n.e1a, nullptr, D, H, W, n.e1b, s1, s, z1)) ||
      (rc = pool(n.e1b, n.p1, D, H, W, s1, s)) ||
      (rc = conv(n.layer[2], n.p1, nullptr, D2, H2, W2, n.e2a, s2, s, z2)) ||
      (rc = conv(n.layer[3], n.e2a, nullptr, D2, H2, W2, n.e2b, s2, s, z2)) ||
      (rc = pool(n.e2b, n.p2, D2, H2, W2, s2, s)) ||
      (rc = conv(n.layer[4], n.p2, nullptr, D4, H4, W4, n.ba, s3, s, z4)) ||
      (rc = conv(n.layer[5], n.ba, nullptr, D4, H4, W4, n.bb, s3, s, z4)) ||
      (rc = conv(n.layer[6], n.e2b, n.bb, D2, H2, W2, n.d2a, s2, s, z2)) ||
      (rc = conv(n.layer[7], n.d2a, nullptr, D2, H2, W2, n.d2b, s2, s, z2)) ||
      (rc = conv(n.layer[8], n.e1b, n.d2b, D, H, W, n.d1a, s1, s, z1)) ||
      (rc = conv(n.layer[9], n.d1a, nullptr, D, H, W, n.d1b, s1, s, z1)))
    return rc;
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

template <typename T>
int
dalloc0(T ** p, size_t n)
{
  *p = nullptr;
  DOMA_HIP(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
  DOMA_HIP(hipMemset(*p, 0, n * sizeof(T)));
  return DOMA_OK;
}
} // namespace

void
net_free(Net & n)
{
  for (auto & L : n.layer)
  {
    dfree(L.w);
    dfree(L.bias);
  }
  dfree(n.feat);
  dfree(n.e1a);
  dfree(n.e1b);
  dfree(n.p1);
  dfree(n.e2a);
  dfree(n.e2b);
  dfree(n.p2);
  dfree(n.ba);
  dfree(n.bb);
  dfree(n.d2a);
  dfree(n.d2b);
  dfree(n.d1a);
  dfree(n.d1b);
  dfree(n.red);
  dfree(n.scale);
  n.ready = false;
}

int
net_setup(Net & n, const float * params, int D, int H, int W)
{
  net_free(n);
  if (D % 4 || H % 4 || W % 4 || D < 4 || H < 4 || W < 4)
    return fail(DOMA_ERR_INVALID, "the learned correction needs every extent divisible by 4");
  n.D = D;
  n.H = H;
  n.W = W;
  const int64_t off_total = total_params();
  (void)off_total;
  int64_t off = 0;
  int rc;
  for (int l = 0; l < kLayers; ++l)
  {
    const LayerSpec & sp = kSpec[l];
    PackedLayer & L = n.layer[l];
    const int cin = sp.cin_a + sp.cin_b;
    const float * Wt = params + off;
    const float * Bs = Wt + (int64_t)sp.cout * sp.taps * cin;
    off += layer_params(sp);
    L.cout = sp.cout;
    L.relu = sp.relu;
    L.taps = sp.taps;
    if (sp.taps == 1)
    {
      // Head: FP32 weights then bias in one array.
      std::vector<float> wb(cin + 1);
      for (int ci = 0; ci < cin; ++ci)
        wb[ci] = Wt[ci];
      wb[cin] = Bs[0];
      if ((rc = dalloc0(&L.bias, wb.size())))
        return rc;
      DOMA_HIP(hipMemcpy(L.bias, wb.data(), wb.size() * 4, hipMemcpyHostToDevice));
      L.cin_a_p = cin;
      continue;
    }
    L.cin_a_p = pad8(sp.cin_a);
    L.cin_b_p = sp.cin_b ? pad8(sp.cin_b) : 0;
    const int cin_p = L.cin_a_p + L.cin_b_p;
    L.cout16 = (sp.cout + 15) / 16 * 16;
    L.kpad = (27 * cin_p + 31) / 32 * 32;
    std::vector<uint16_t> w((size_t)L.cout16 * L.kpad, 0);
    for (int co = 0; co < sp.cout; ++co)
      for (int t = 0; t < 27; ++t)
        for (int ci = 0; ci < cin; ++ci)
        {
          const int slot = ci < sp.cin_a ? ci : L.cin_a_p + (ci - sp.cin_a);
          w[(size_t)co * L.kpad + t * cin_p + slot] = f2bf(Wt[((size_t)co * 27 + t) * cin + ci]);
        }
    if ((rc = dalloc0(&L.w, w.size())) || (rc = dalloc0(&L.bias, (size_t)sp.cout)))
      return rc;
    DOMA_HIP(hipMemcpy(L.w, w.data(), w.size() * 2, hipMemcpyHostToDevice));
    DOMA_HIP(hipMemcpy(L.bias, Bs, sp.cout * 4, hipMemcpyHostToDevice));
  }
  const size_t V = (size_t)D * H * W, V2 = V / 8, V4 = V / 64;
  // Zero-filled once: padding channels are never written and must stay finite.
  if ((rc = dalloc0(&n.feat, V * 8)) || (rc = dalloc0(&n.e1a, V * pad8(C1))) ||
      (rc = dalloc0(&n.e1b, V * pad8(C1))) || (rc = dalloc0(&n.p1, V2 * pad8(C1))) ||
      (rc = dalloc0(&n.e2a, V2 * pad8(C2))) || (rc = dalloc0(&n.e2b, V2 * pad8(C2))) ||
      (rc = dalloc0(&n.p2, V4 * pad8(C2))) || (rc = dalloc0(&n.ba, V4 * pad8(C3))) ||
      (rc = dalloc0(&n.bb, V4 * pad8(C3))) || (rc = dalloc0(&n.d2a, V2 * pad8(C2))) ||
      (rc = dalloc0(&n.d2b, V2 * pad8(C2))) || (rc = dalloc0(&n.d1a, V * pad8(C1))) ||
      (rc = dalloc0(&n.d1b, V * pad8(C1))) || (rc = dalloc0(&n.red, 1024)) ||
      (rc = dalloc0(&n.scale, 2)))
    return rc;
  n.ready = true;
  return DOMA_OK;
}

int
net_apply(Net & n, const double * dlast, const double * dprev, const int32_t * mat,
          const double * sigt, const double * sigs, int G, int g, double * corr, hipStream_t s)
{
  if (!n.ready)
    return fail(DOMA_ERR_STATE, "the learned correction has no weights");
  const int N = n.D * n.H * n.W;
  const int nb = 512;  // two sets of partial sums in n.red (1024 doubles)
  hipLaunchKernelGGL(sumsq_kernel, dim3(nb), dim3(RT), 0, s, dlast, N, n.red);
  hipLaunchKernelGGL(sumsq_kernel, dim3(nb), dim3(RT), 0, s, dprev, N, n.red + nb);
  hipLaunchKernelGGL(rms_final_kernel, dim3(1), dim3(RT), 0, s, n.red, nb, N, n.scale);
  hipLaunchKernelGGL(features_kernel, dim3(nblk(N, 256)), dim3(256), 0, s, dlast, dprev, n.scale,
                     mat, sigt, sigs, G, g, N, n.feat);
  int rc = body(n, s);
  if (rc)
    return rc;
  hipLaunchKernelGGL(head_corr_kernel, dim3(nblk(N, 256)), dim3(256), 0, s, n.d1b, pad8(C1), C1,
                     n.layer[10].bias, n.scale, dlast, N, corr);
  return DOMA_OK;
}

int
net_forward_features(Net & n, const float * feat_dev, float * out_dev, hipStream_t s)
{
  if (!n.ready)
    return fail(DOMA_ERR_STATE, "the learned correction has no weights");
  const int N = n.D * n.H * n.W;
  hipLaunchKernelGGL(features_f32_kernel, dim3(nblk(N, 256)), dim3(256), 0, s, feat_dev, N,
                     n.feat);
  int rc = body(n, s);
  if (rc)
    return rc;
  hipLaunchKernelGGL(head_raw_kernel, dim3(nblk(N, 256)), dim3(256), 0, s, n.d1b, pad8(C1), C1,
                     n.layer[10].bias, N, out_dev);
  return DOMA_OK;
}

// ======================================================================== //
// Training: forward on a batch of cubic crops stacked along z, hand-written //
// backward (data gradients on the same MFMA convolution with flipped and    //
// transposed weights, weight gradients by a deterministic chunked FP32      //
// reduction), relative-L2 loss, global-norm clipping and Adam on FP32       //
// master weights.                                                           //
// ======================================================================== //
namespace
{
constexpr int WG_CHUNK = 1024;  // voxels per weight-gradient partial

// Slot of input channel ci of a layer in its padded concat layout.
__device__ inline int
in_slot(int ci, int cin_a, int ca_p)
{
  return ci < cin_a ? ci : ca_p + (ci - cin_a);
}

// FP32 layer weights W[cout][27][cin] -> forward BF16 [cout16][kpad], K = tap cin_p + slot.
__global__ void
pack_fwd_kernel(const float * W, int cout, int cin_a, int cin_b, int ca_p, int cb_p, int kpad,
                int cout16, uint16_t * out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cout16 * kpad)
    return;
  const int co = i / kpad, k = i - co * kpad;
  const int cin_p = ca_p + cb_p, cin = cin_a + cin_b;
  const int tap = k / cin_p, slot = k - tap * cin_p;
  int ci = -1;
  if (slot < ca_p)
    ci = slot < cin_a ? slot : -1;
  else
    ci = slot - ca_p < cin_b ? cin_a + (slot - ca_p) : -1;
  float v = 0.f;
  if (co < cout && tap < 27 && ci >= 0)
    v = W[((size_t)co * 27 + tap) * cin + ci];
  out[i] = f2bf(v);
}

// Data-gradient weights: rows = the forward layer's padded input slots, K =
// tap' cout_p + co with tap' = 26 - tap (the stencil mirrored).
__global__ void
pack_bwd_kernel(const float * W, int cout, int cin_a, int cin_b, int ca_p, int cb_p, int cout_p,
                int kpad, int rows16, uint16_t * out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows16 * kpad)
    return;
  const int r = i / kpad, k = i - r * kpad;
  const int cin = cin_a + cin_b;
  const int tapp = k / cout_p, co = k - tapp * cout_p;
  int ci = -1;
  if (r < ca_p)
    ci = r < cin_a ? r : -1;
  else if (r < ca_p + cb_p)
    ci = r - ca_p < cin_b ? cin_a + (r - ca_p) : -1;
  float v = 0.f;
  if (tapp < 27 && co < cout && ci >= 0)
    v = W[((size_t)co * 27 + (26 - tapp)) * cin + ci];
  out[i] = f2bf(v);
}

__global__ void
copy_f32_kernel(const float * in, int n, float * out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = in[i];
}

struct WgArgs
{
  const void * dy;     // BF16 [V][dy_stride] or FP32 [V] (head)
  int dy_stride;
  int dy_f32;
  const uint16_t * a;  // forward input, full resolution [V][ca_p]
  int ca_p;
  const uint16_t * b;  // forward input, coarse [V/8][cb_p] (or null)
  int cb_p;
  int cin_a, cin_b, cout, taps;
  int D, H, W, V, Ds;
  int nchunks, nj;
};

// partial[chunk][j], j = flat index of (co, tap, ci) then the biases.
__global__ __launch_bounds__(256) void
wgrad_kernel(WgArgs p, float * partial)
{
  const int j = blockIdx.y * 256 + threadIdx.x;
  const int chunk = blockIdx.x;
  if (j >= p.nj)
    return;
  const int cin = p.cin_a + p.cin_b;
  const int nw = p.cout * p.taps * cin;
  int co, tap = -1, ci = 0;
  if (j < nw)
  {
    co = j / (p.taps * cin);
    const int r = j - co * p.taps * cin;
    tap = r / cin;
    ci = r - tap * cin;
  }
  else
    co = j - nw;
  const int slot = in_slot(ci, p.cin_a, p.ca_p);
  int dz = 0, dy = 0, dx = 0;
  if (p.taps == 27 && tap >= 0)
  {
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
  }
  const int H2 = p.H >> 1, W2 = p.W >> 1;
  const int v0 = chunk * WG_CHUNK, v1 = min(p.V, v0 + WG_CHUNK);
  float acc = 0.f;
  for (int v = v0; v < v1; ++v)
  {
    const float g = p.dy_f32 ? reinterpret_cast<const float *>(p.dy)[v]
                             : bf2f(reinterpret_cast<const uint16_t *>(p.dy)[(size_t)v * p.dy_stride + co]);
    float x = 1.f;
    if (tap >= 0)
    {
      const int vx = v % p.W, vy = (v / p.W) % p.H, vz = v / (p.W * p.H);
      const int zl = vz % p.Ds + dz, z = vz + dz, y = vy + dy, xx = vx + dx;
      x = 0.f;
      if (zl >= 0 && zl < p.Ds && y >= 0 && y < p.H && xx >= 0 && xx < p.W)
      {
        if (slot < p.ca_p)
          x = bf2f(p.a[((size_t)(z * p.H + y) * p.W + xx) * p.ca_p + slot]);
        else
          x = bf2f(p.b[((size_t)((z >> 1) * H2 + (y >> 1)) * W2 + (xx >> 1)) * p.cb_p +
                       (slot - p.ca_p)]);
      }
    }
    acc = fmaf(g, x, acc);
  }
  partial[(size_t)chunk * p.nj + j] = acc;
}

__global__ void
wgrad_final_kernel(const float * partial, int nchunks, int nj, float * grad)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj)
    return;
  float s = 0.f;
  for (int c = 0; c < nchunks; ++c)
    s += partial[(size_t)c * nj + j];
  grad[j] = s;
}

// Loss terms: block partials of sum (p - y)^2 and sum y^2.
__global__ __launch_bounds__(256) void
loss_partial_kernel(const float * pred, const float * y, int V, float * red)
{
  __shared__ float s1[256], s2[256];
  float a = 0.f, b = 0.f;
  for (int v = blockIdx.x * 256 + threadIdx.x; v < V; v += gridDim.x * 256)
  {
    const float d = pred[v] - y[v];
    a += d * d;
    b += y[v] * y[v];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1)
  {
    if ((int)threadIdx.x < k)
    {
      s1[threadIdx.x] += s1[threadIdx.x + k];
      s2[threadIdx.x] += s2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
  {
    red[2 * blockIdx.x] = s1[0];
    red[2 * blockIdx.x + 1] = s2[0];
  }
}

// out[0] = loss = sum (p-y)^2 / (sum y^2 + eps), out[1] = the denominator.
__global__ __launch_bounds__(256) void
loss_final_kernel(const float * red, int nb, float * out)
{
  __shared__ float s1[256], s2[256];
  float a = 0.f, b = 0.f;
  for (int k = threadIdx.x; k < nb; k += 256)
  {
    a += red[2 * k];
    b += red[2 * k + 1];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1)
  {
    if ((int)threadIdx.x < k)
    {
      s1[threadIdx.x] += s1[threadIdx.x + k];
      s2[threadIdx.x] += s2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
  {
    const float den = s2[0] + 1e-12f;
    out[0] = s1[0] / den;
    out[1] = den;
  }
}

__global__ void
dpred_kernel(const float * pred, const float * y, const float * lossv, int V, float * dpred)
{
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V)
    dpred[v] = 2.f * (pred[v] - y[v]) / lossv[1];
}

// Gradient at the head's input, masked by the ReLU of the last hidden layer.
__global__ void
head_back_kernel(const float * dpred, const float * wh, const uint16_t * h, int C, int stride,
                 int V, uint16_t * dy)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)V * stride)
    return;
  const int c = (int)(i % stride);
  const size_t v = i / stride;
  float g = 0.f;
  if (c < C && bf2f(h[i]) > 0.f)
    g = dpred[v] * wh[c];
  dy[i] = f2bf(g);
}

// dy[v][c] = src[v][c] if act[v][c] > 0 else 0 (c < C; padded channels 0).
template <typename T>
__global__ void
mask_kernel(const T * src, int src_stride, const uint16_t * act, int act_stride, int C, int V,
            uint16_t * dy)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)V * act_stride)
    return;
  const int c = (int)(i % act_stride);
  const size_t v = i / act_stride;
  float g = 0.f;
  if (c < C && bf2f(act[i]) > 0.f)
  {
    if constexpr (sizeof(T) == 2)
      g = bf2f(src[v * src_stride + c]);
    else
      g = src[v * src_stride + c];
  }
  dy[i] = f2bf(g);
}

// Skip part of a concat data gradient: acc[v][c] = g[v][c] (c < C).
__global__ void
split_skip_kernel(const uint16_t * g, int g_stride, int C, int acc_stride, int V, float * acc)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)V * acc_stride)
    return;
  const int c = (int)(i % acc_stride);
  const size_t v = i / acc_stride;
  acc[i] = c < C ? bf2f(g[v * g_stride + c]) : 0.f;
}

// Upsampling backward: low[u][c] = sum over the 8 children of g[child][off + c].
__global__ void
up_back_kernel(const uint16_t * g, int g_stride, int off, int C, int low_stride, int D, int H,
               int W, float * low)
{
  const int D2 = D >> 1, H2 = H >> 1, W2 = W >> 1;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)D2 * H2 * W2 * low_stride)
    return;
  const int c = (int)(i % low_stride);
  const size_t u = i / low_stride;
  float s = 0.f;
  if (c < C)
  {
    const int x = (int)(u % W2), y = (int)((u / W2) % H2), z = (int)(u / ((size_t)W2 * H2));
    for (int d = 0; d < 8; ++d)
    {
      const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
      s += bf2f(g[((size_t)(zz * H + yy) * W + xx) * g_stride + off + c]);
    }
  }
  low[i] = s;
}

// Average-pooling backward, accumulated: acc[child][c] += g[parent][c] / 8.
__global__ void
pool_back_add_kernel(const uint16_t * g, int g_stride, int C, int acc_stride, int D, int H, int W,
                     float * acc)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)D * H * W * acc_stride)
    return;
  const int c = (int)(i % acc_stride);
  if (c >= C)
    return;
  const size_t v = i / acc_stride;
  const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((size_t)W * H));
  const size_t u = ((size_t)(z >> 1) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1);
  acc[i] += 0.125f * bf2f(g[u * g_stride + c]);
}

__global__ __launch_bounds__(256) void
sumsq_f32_kernel(const float * g, int n, float * red)
{
  __shared__ float sh[256];
  float a = 0.f;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    a += g[i] * g[i];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1)
  {
    if ((int)threadIdx.x < k)
      sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    red[blockIdx.x] = sh[0];
}

// Adam with global-norm clipping (max norm 1): red holds nb partial squared norms.
__global__ void
adam_kernel(float * w, const float * g, float * m, float * v, int n, const float * red, int nb,
            float lr, float b1, float b2, float eps, int step)
{
  __shared__ float scale;
  if (threadIdx.x == 0)
  {
    float s = 0.f;
    for (int k = 0; k < nb; ++k)
      s += red[k];
    const float norm = sqrtf(s);
    scale = norm > 1.f ? 1.f / (norm + 1e-6f) : 1.f;
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float gi = g[i] * scale;
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, (float)step));
  const float vh = vi / (1.f - powf(b2, (float)step));
  w[i] -= lr * mh / (sqrtf(vh) + eps);
}
} // namespace

struct Trainer
{
  int C = 0, nb = 0;
  Net net;
  float * params = nullptr;  // FP32 master weights [P]
  float * grads = nullptr;
  float * m = nullptr;
  float * v = nullptr;
  PackedLayer bwd[kLayers];  // data-gradient weights (layers 1-9)
  float * zero = nullptr;     // [64]
  float * fin = nullptr;      // [V][CIN]
  float * y = nullptr;        // [V]
  float * pred = nullptr;     // [V]
  float * dpred = nullptr;    // [V]
  uint16_t * dyA = nullptr;   // [V][64]
  uint16_t * dyB = nullptr;   // [V][64]
  uint16_t * gtmp = nullptr;  // [V][64]
  float * acc1 = nullptr;     // [V][16]
  float * acc2 = nullptr;     // [V/8][24]
  float * low = nullptr;      // [V/8][32]
  float * partial = nullptr;
  size_t partial_cap = 0;
  float * red = nullptr;      // [1024]
  float * lossv = nullptr;    // [2]
  int64_t off[kLayers]{};     // flat offset of each layer
  int step = 0;
};

namespace
{
template <typename T>
int
alloc0(T ** p, size_t n)
{
  *p = nullptr;
  DOMA_HIP(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
  DOMA_HIP(hipMemset(*p, 0, n * sizeof(T)));
  return DOMA_OK;
}

int
repack(Trainer & t, hipStream_t s)
{
  for (int l = 0; l < kLayers; ++l)
  {
    const LayerSpec & sp = kSpec[l];
    PackedLayer & L = t.net.layer[l];
    const float * W = t.params + t.off[l];
    const int cin = sp.cin_a + sp.cin_b;
    if (sp.taps == 1)
    {
      hipLaunchKernelGGL(copy_f32_kernel, dim3(1), dim3(64), 0, s, W, cin + 1, L.bias);
      continue;
    }
    hipLaunchKernelGGL(pack_fwd_kernel, dim3(nblk((size_t)L.cout16 * L.kpad, 256)), dim3(256), 0,
                       s, W, sp.cout, sp.cin_a, sp.cin_b, L.cin_a_p, L.cin_b_p, L.kpad, L.cout16,
                       L.w);
    hipLaunchKernelGGL(copy_f32_kernel, dim3(1), dim3(64), 0, s,
                       W + (int64_t)sp.cout * 27 * cin, sp.cout, L.bias);
    if (l == 0)
      continue;
    PackedLayer & B = t.bwd[l];
    hipLaunchKernelGGL(pack_bwd_kernel, dim3(nblk((size_t)B.cout16 * B.kpad, 256)), dim3(256), 0,
                       s, W, sp.cout, sp.cin_a, sp.cin_b, L.cin_a_p, L.cin_b_p, B.cin_a_p, B.kpad,
                       B.cout16, B.w);
  }
  return DOMA_OK;
}

// Weight and bias gradients of layer l into t.grads.
int
wgrad(Trainer & t, int l, const void * dy, int dy_stride, bool dy_f32, const uint16_t * a,
      const uint16_t * b, int D, int H, int W, int Ds, hipStream_t s)
{
  const LayerSpec & sp = kSpec[l];
  const PackedLayer & L = t.net.layer[l];
  WgArgs p{};
  p.dy = dy;
  p.dy_stride = dy_stride;
  p.dy_f32 = dy_f32 ? 1 : 0;
  p.a = a;
  p.ca_p = sp.taps == 1 ? pad8(sp.cin_a) : L.cin_a_p;
  p.b = b;
  p.cb_p = L.cin_b_p;
  p.cin_a = sp.cin_a;
  p.cin_b = sp.cin_b;
  p.cout = sp.cout;
  p.taps = sp.taps;
  p.D = D;
  p.H = H;
  p.W = W;
  p.V = D * H * W;
  p.Ds = Ds;
  p.nchunks = (p.V + WG_CHUNK - 1) / WG_CHUNK;
  p.nj = (int)layer_params(sp);
  const size_t need = (size_t)p.nchunks * p.nj;
  if (need > t.partial_cap)
    return fail(DOMA_ERR_INVALID, "weight-gradient workspace too small");
  hipLaunchKernelGGL(wgrad_kernel, dim3(p.nchunks, nblk(p.nj, 256)), dim3(256), 0, s, p,
                     t.partial);
  hipLaunchKernelGGL(wgrad_final_kernel, dim3(nblk(p.nj, 256)), dim3(256), 0, s, t.partial,
                     p.nchunks, p.nj, t.grads + t.off[l]);
  return DOMA_OK;
}

// Data gradient of layer l: out[v][slot] over the forward input's padded slots.
int
dgrad(Trainer & t, int l, const uint16_t * dy, int D, int H, int W, int Ds, uint16_t * out,
      int out_stride, hipStream_t s)
{
  const PackedLayer & B = t.bwd[l];
  const PackedLayer & L = t.net.layer[l];
  ConvArgs p{};
  p.a = dy;
  p.ca = B.cin_a_p;
  p.b = nullptr;
  p.cb = 0;
  p.D = D;
  p.H = H;
  p.W = W;
  p.V = D * H * W;
  p.Ds = Ds;
  p.w = B.w;
  p.bias = t.zero;
  p.cout = L.cin_a_p + L.cin_b_p;
  p.kpad = B.kpad;
  p.cin_p = B.cin_a_p;
  p.out = out;
  p.out_stride = out_stride;
  p.relu = 0;
  return launch_conv(p, B.cout16, s);
}

template <typename T>
int
mask(const T * src, int src_stride, const uint16_t * act, int act_stride, int C, int V,
     uint16_t * dy, hipStream_t s)
{
  hipLaunchKernelGGL((mask_kernel<T>), dim3(nblk((size_t)V * act_stride, 256)), dim3(256), 0, s,
                     src, src_stride, act, act_stride, C, V, dy);
  return DOMA_OK;
}

int
forward(Trainer & t, hipStream_t s)
{
  Net & n = t.net;
  const int V = n.D * n.H * n.W;
  hipLaunchKernelGGL(features_f32_kernel, dim3(nblk(V, 256)), dim3(256), 0, s, t.fin, V, n.feat);
  int rc = body(n, s, t.C);
  if (rc)
    return rc;
  hipLaunchKernelGGL(head_raw_kernel, dim3(nblk(V, 256)), dim3(256), 0, s, n.d1b, pad8(C1), C1,
                     n.layer[10].bias, V, t.pred);
  return DOMA_OK;
}

int
backward(Trainer & t, hipStream_t s)
{
  Net & n = t.net;
  const int D = n.D, H = n.H, W = n.W, V = D * H * W;
  const int D2 = D / 2, H2 = H / 2, W2 = W / 2, V2 = V / 8, D4 = D / 4, H4 = H / 4, W4 = W / 4;
  const int V4 = V / 64;
  const int z1 = t.C, z2 = t.C / 2, z4 = t.C / 4;
  const int s1 = pad8(C1), s2 = pad8(C2), s3 = pad8(C3);
  int rc;
  hipLaunchKernelGGL(loss_partial_kernel, dim3(256), dim3(256), 0, s, t.pred, t.y, V, t.red);
  hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(256), 0, s, t.red, 256, t.lossv);
  hipLaunchKernelGGL(dpred_kernel, dim3(nblk(V, 256)), dim3(256), 0, s, t.pred, t.y, t.lossv, V,
                     t.dpred);
  // head (1x1x1)
  if ((rc = wgrad(t, 10, t.dpred, 1, true, n.d1b, nullptr, D, H, W, z1, s)))
    return rc;
  hipLaunchKernelGGL(head_back_kernel, dim3(nblk((size_t)V * s1, 256)), dim3(256), 0, s, t.dpred,
                     t.params + t.off[10], n.d1b, C1, s1, V, t.dyA);
  // dec1b
  if ((rc = wgrad(t, 9, t.dyA, s1, false, n.d1a, nullptr, D, H, W, z1, s)) ||
      (rc = dgrad(t, 9, t.dyA, D, H, W, z1, t.gtmp, s1, s)) ||
      (rc = mask(t.gtmp, s1, n.d1a, s1, C1, V, t.dyB, s)))
    return rc;
  // dec1a: concat(e1b, up(d2b))
  const int c8 = s1 + s2;  // 16 + 24 padded input slots
  if ((rc = wgrad(t, 8, t.dyB, s1, false, n.e1b, n.d2b, D, H, W, z1, s)) ||
      (rc = dgrad(t, 8, t.dyB, D, H, W, z1, t.gtmp, c8, s)))
    return rc;
  hipLaunchKernelGGL(split_skip_kernel, dim3(nblk((size_t)V * s1, 256)), dim3(256), 0, s, t.gtmp,
                     c8, C1, s1, V, t.acc1);
  hipLaunchKernelGGL(up_back_kernel, dim3(nblk((size_t)V2 * s2, 256)), dim3(256), 0, s, t.gtmp, c8,
                     s1, C2, s2, D, H, W, t.low);
  if ((rc = mask(t.low, s2, n.d2b, s2, C2, V2, t.dyA, s)))
    return rc;
  // dec2b
  if ((rc = wgrad(t, 7, t.dyA, s2, false, n.d2a, nullptr, D2, H2, W2, z2, s)) ||
      (rc = dgrad(t, 7, t.dyA, D2, H2, W2, z2, t.gtmp, s2, s)) ||
      (rc = mask(t.gtmp, s2, n.d2a, s2, C2, V2, t.dyB, s)))
    return rc;
  // dec2a: concat(e2b, up(bb))
  const int c6 = s2 + s3;  // 24 + 32
  if ((rc = wgrad(t, 6, t.dyB, s2, false, n.e2b, n.bb, D2, H2, W2, z2, s)) ||
      (rc = dgrad(t, 6, t.dyB, D2, H2, W2, z2, t.gtmp, c6, s)))
    return rc;
  hipLaunchKernelGGL(split_skip_kernel, dim3(nblk((size_t)V2 * s2, 256)), dim3(256), 0, s, t.gtmp,
                     c6, C2, s2, V2, t.acc2);
  hipLaunchKernelGGL(up_back_kernel, dim3(nblk((size_t)V4 * s3, 256)), dim3(256), 0, s, t.gtmp, c6,
                     s2, C3, s3, D2, H2, W2, t.low);
  if ((rc = mask(t.low, s3, n.bb, s3, C3, V4, t.dyA, s)))
    return rc;
  // bottleneck
  if ((rc = wgrad(t, 5, t.dyA, s3, false, n.ba, nullptr, D4, H4, W4, z4, s)) ||
      (rc = dgrad(t, 5, t.dyA, D4, H4, W4, z4, t.gtmp, s3, s)) ||
      (rc = mask(t.gtmp, s3, n.ba, s3, C3, V4, t.dyB, s)) ||
      (rc = wgrad(t, 4, t.dyB, s3, false, n.p2, nullptr, D4, H4, W4, z4, s)) ||
      (rc = dgrad(t, 4, t.dyB, D4, H4, W4, z4, t.gtmp, s2, s)))
    return rc;
  hipLaunchKernelGGL(pool_back_add_kernel, dim3(nblk((size_t)V2 * s2, 256)), dim3(256), 0, s,
                     t.gtmp, s2, C2, s2, D2, H2, W2, t.acc2);
  if ((rc = mask(t.acc2, s2, n.e2b, s2, C2, V2, t.dyA, s)))
    return rc;
  // encoder level 2
  if ((rc = wgrad(t, 3, t.dyA, s2, false, n.e2a, nullptr, D2, H2, W2, z2, s)) ||
      (rc = dgrad(t, 3, t.dyA, D2, H2, W2, z2, t.gtmp, s2, s)) ||
      (rc = mask(t.gtmp, s2, n.e2a, s2, C2, V2, t.dyB, s)) ||
      (rc = wgrad(t, 2, t.dyB, s2, false, n.p1, nullptr, D2, H2, W2, z2, s)) ||
      (rc = dgrad(t, 2, t.dyB, D2, H2, W2, z2, t.gtmp, s1, s)))
    return rc;
  hipLaunchKernelGGL(pool_back_add_kernel, dim3(nblk((size_t)V * s1, 256)), dim3(256), 0, s,
                     t.gtmp, s1, C1, s1, D, H, W, t.acc1);
  if ((rc = mask(t.acc1, s1, n.e1b, s1, C1, V, t.dyA, s)))
    return rc;
  // encoder level 1
  if ((rc = wgrad(t, 1, t.dyA, s1, false, n.e1a, nullptr, D, H, W, z1, s)) ||
      (rc = dgrad(t, 1, t.dyA, D, H, W, z1, t.gtmp, s1, s)) ||
      (rc = mask(t.gtmp, s1, n.e1a, s1, C1, V, t.dyB, s)) ||
      (rc = wgrad(t, 0, t.dyB, s1, false, n.feat, nullptr, D, H, W, z1, s)))
    return rc;
  return DOMA_OK;
}
} // namespace

void
trainer_free(Trainer * t)
{
  if (!t)
    return;
  net_free(t->net);
  for (auto & B : t->bwd)
  {
    dfree(B.w);
    dfree(B.bias);
  }
  dfree(t->params);
  dfree(t->grads);
  dfree(t->m);
  dfree(t->v);
  dfree(t->zero);
  dfree(t->fin);
  dfree(t->y);
  dfree(t->pred);
  dfree(t->dpred);
  dfree(t->dyA);
  dfree(t->dyB);
  dfree(t->gtmp);
  dfree(t->acc1);
  dfree(t->acc2);
  dfree(t->low);
  dfree(t->partial);
  dfree(t->red);
  dfree(t->lossv);
  delete t;
}

int
trainer_create(int crop, int batch, const float * params, Trainer ** out)
{
  *out = nullptr;
  if (crop < 4 || crop % 4 || batch < 1)
    return fail(DOMA_ERR_INVALID, "trainer: crop must be a multiple of 4 and batch >= 1");
  Trainer * t = new (std::nothrow) Trainer();
  if (!t)
    return fail(DOMA_ERR_INVALID, "out of host memory");
  t->C = crop;
  t->nb = batch;
  int rc = net_setup(t->net, params, batch * crop, crop, crop);
  const int64_t P = total_params();
  const size_t V = (size_t)batch * crop * crop * crop;
  int64_t o = 0;
  for (int l = 0; l < kLayers; ++l)
  {
    t->off[l] = o;
    o += layer_params(kSpec[l]);
  }
  size_t cap = 0;
  for (int l = 0; l < kLayers; ++l)
  {
    const size_t vl = l <= 1 || l >= 8 ? V : (l == 4 || l == 5 ? V / 64 : V / 8);
    cap = std::max(cap, ((vl + WG_CHUNK - 1) / WG_CHUNK) * (size_t)layer_params(kSpec[l]));
  }
  if (!rc)
  {
    if ((rc = alloc0(&t->params, P)) || (rc = alloc0(&t->grads, P)) || (rc = alloc0(&t->m, P)) ||
        (rc = alloc0(&t->v, P)) || (rc = alloc0(&t->zero, 64)) || (rc = alloc0(&t->fin, V * CIN)) ||
        (rc = alloc0(&t->y, V)) || (rc = alloc0(&t->pred, V)) || (rc = alloc0(&t->dpred, V)) ||
        (rc = alloc0(&t->dyA, V * 64)) || (rc = alloc0(&t->dyB, V * 64)) ||
        (rc = alloc0(&t->gtmp, V * 64)) || (rc = alloc0(&t->acc1, V * 16)) ||
        (rc = alloc0(&t->acc2, V / 8 * 24)) || (rc = alloc0(&t->low, V / 8 * 32)) ||
        (rc = alloc0(&t->partial, cap)) || (rc = alloc0(&t->red, 1024)) ||
        (rc = alloc0(&t->lossv, 2)))
      ;
    t->partial_cap = cap;
  }
  for (int l = 1; !rc && l < kLayers - 1; ++l)
  {
    const LayerSpec & sp = kSpec[l];
    const PackedLayer & L = t->net.layer[l];
    PackedLayer & B = t->bwd[l];
    B.cin_a_p = pad8(sp.cout);
    B.cin_b_p = 0;
    const int rows = L.cin_a_p + L.cin_b_p;
    B.cout = rows;
    B.cout16 = rows <= 16 ? 16 : (rows <= 32 ? 32 : 64);
    B.kpad = (27 * B.cin_a_p + 31) / 32 * 32;
    rc = alloc0(&B.w, (size_t)B.cout16 * B.kpad);
  }
  if (!rc)
  {
    if (hipMemcpy(t->params, params, P * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(DOMA_ERR_DEVICE, "trainer: upload failed");
  }
  if (!rc)
    rc = repack(*t, nullptr);
  if (!rc && hipDeviceSynchronize() != hipSuccess)
    rc = fail(DOMA_ERR_DEVICE, "trainer: set-up failed");
  if (rc)
  {
    trainer_free(t);
    return rc;
  }
  *out = t;
  return DOMA_OK;
}

int
trainer_run(Trainer * t, const float * feats, const float * targets, float lr, bool update,
            float * loss, float * grad_out)
{
  const size_t V = (size_t)t->nb * t->C * t->C * t->C;
  hipStream_t s = nullptr;
  DOMA_HIP(hipMemcpy(t->fin, feats, V * CIN * sizeof(float), hipMemcpyHostToDevice));
  DOMA_HIP(hipMemcpy(t->y, targets, V * sizeof(float), hipMemcpyHostToDevice));
  int rc;
  if ((rc = forward(*t, s)) || (rc = backward(*t, s)))
    return rc;
  const int P = (int)total_params();
  if (update)
  {
    ++t->step;
    hipLaunchKernelGGL(sumsq_f32_kernel, dim3(256), dim3(256), 0, s, t->grads, P, t->red);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(P, 256)), dim3(256), 0, s, t->params, t->grads, t->m,
                       t->v, P, t->red, 256, lr, 0.9f, 0.999f, 1e-8f, t->step);
    if ((rc = repack(*t, s)))
      return rc;
  }
  float lv[2];
  DOMA_HIP(hipMemcpy(lv, t->lossv, 2 * sizeof(float), hipMemcpyDeviceToHost));
  if (loss)
    *loss = lv[0];
  if (grad_out)
    DOMA_HIP(hipMemcpy(grad_out, t->grads, (size_t)P * sizeof(float), hipMemcpyDeviceToHost));
  DOMA_HIP(hipDeviceSynchronize());
  DOMA_HIP(hipGetLastError());
  return DOMA_OK;
}

int
trainer_params(Trainer * t, float * params)
{
  DOMA_HIP(hipMemcpy(params, t->params, total_params() * sizeof(float), hipMemcpyDeviceToHost));
  return DOMA_OK;
}

// ------------------------------------------------------------------------- //
// Forward pass without allocation                                            //
// ------------------------------------------------------------------------- //
//
// doma_net_forward_noalloc runs the same layers as doma_net_forward on device
// buffers the caller owns, with the packed BF16 weights and every activation in
// one file-scope device array (`scratch`), sized for grids of up to
// DOMA_NET_NOALLOC_MAX_VOXELS voxels.  Host code never uses the array's name as
// an address: its device address is looked up once per device with
// hipGetSymbolAddress and kept in noalloc_base[].
namespace
{
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Byte offsets of the scratch regions (weights first, then activations sized for
// the largest grid); every region starts on a 256-byte boundary.
struct NoAllocLayout
{
  size_t w[kLayers] = {};  // packed weights of the 3x3x3 layers
  size_t feat = 0, e1a = 0, e1b = 0, p1 = 0, e2a = 0, e2b = 0, p2 = 0, ba = 0, bb = 0, d2a = 0,
         d2b = 0, d1a = 0, d1b = 0, act_begin = 0, bytes = 0;
};

constexpr int64_t layer_kpad(const LayerSpec & sp)
{
  return (27 * (pad8(sp.cin_a) + (sp.cin_b ? pad8(sp.cin_b) : 0)) + 31) / 32 * 32;
}

constexpr NoAllocLayout noalloc_layout()
{
  NoAllocLayout L{};
  size_t off = 0;
  for (int l = 0; l < kLayers; ++l)
  {
    const LayerSpec & sp = kSpec[l];
    if (sp.taps != 27)
      continue;
    L.w[l] = off;
    off = align_up(off + (size_t)((sp.cout + 15) / 16 * 16) * layer_kpad(sp) * 2);
  }
  const size_t V = DOMA_NET_NOALLOC_MAX_VOXELS, V2 = V / 8, V4 = V / 64;
  L.act_begin = off;
  auto take = [&off](size_t elems) {
    const size_t at = off;
    off = align_up(off + elems * 2);
    return at;
  };
  L.feat = take(V * 8);
  L.e1a = take(V * pad8(C1));
  L.e1b = take(V * pad8(C1));
  L.p1 = take(V2 * pad8(C1));
  L.e2a = take(V2 * pad8(C2));
  L.e2b = take(V2 * pad8(C2));
  L.p2 = take(V4 * pad8(C2));
  L.ba = take(V4 * pad8(C3));
  L.bb = take(V4 * pad8(C3));
  L.d2a = take(V2 * pad8(C2));
  L.d2b = take(V2 * pad8(C2));
  L.d1a = take(V * pad8(C1));
  L.d1b = take(V * pad8(C1));
  L.bytes = off;
  return L;
}
constexpr NoAllocLayout kNoAlloc = noalloc_layout();
static_assert(kNoAlloc.bytes <= ((size_t)8 << 20), "the forward-pass scratch stays under 8 MiB");

// The scratch array itself (one copy per device, alive until the process ends).
static __device__ __align__(256) unsigned char scratch[kNoAlloc.bytes];

constexpr int kMaxDevices = 64;
unsigned char * noalloc_base[kMaxDevices] = {};
std::mutex noalloc_mu;

// Device address of `scratch` on the current device (looked up on the first
// call per device, with that device current).
int noalloc_scratch(unsigned char ** out)
{
  int dev = 0;
  DOMA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(DOMA_ERR_DEVICE, "doma_net_forward_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_base[dev])
  {
    void * p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DOMA_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                       hipGetErrorString(e));
    noalloc_base[dev] = static_cast<unsigned char *>(p);
  }
  *out = noalloc_base[dev];
  return DOMA_OK;
}

// W[co][tap][cin] (FP32, flat parameter layout) -> BF16 [cout16][kpad] with
// K = tap * cin_p + slot, zero in the padding slots (the layout net_setup packs).
__global__ void
pack_layer_kernel(const float * Wt, int cout, int cin_a, int cin_b, int cin_a_p, int cin_p,
                  int kpad, int n, uint16_t * w)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int co = i / kpad, k = i - co * kpad;
  const int tap = k / cin_p, slot = k - tap * cin_p;
  int ci = -1;
  if (slot < cin_a_p)
    ci = slot < cin_a ? slot : -1;
  else if (slot - cin_a_p < cin_b)
    ci = cin_a + (slot - cin_a_p);
  const int cin = cin_a + cin_b;
  w[i] = (co < cout && tap < 27 && ci >= 0) ? f2bf(Wt[((size_t)co * 27 + tap) * cin + ci]) : 0;
}

__global__ void
fill_u32_kernel(uint32_t * p, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = 0u;
}
} // namespace

int
net_forward_noalloc(const float * d_params, int D, int H, int W, const float * d_features,
                    float * d_out, hipStream_t s)
{
  const size_t V = (size_t)D * H * W;
  if (D % 4 || H % 4 || W % 4 || D < 4 || H < 4 || W < 4)
    return fail(DOMA_ERR_INVALID, "doma_net_forward_noalloc: every extent must be a positive "
                                  "multiple of 4");
  if (V > DOMA_NET_NOALLOC_MAX_VOXELS)
    return fail(DOMA_ERR_INVALID, "doma_net_forward_noalloc: " + std::to_string(V) +
                                      " voxels exceed the scratch limit of " +
                                      std::to_string(DOMA_NET_NOALLOC_MAX_VOXELS));
  unsigned char * base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc)
    return rc;
  // Activations: padding channels are read (with zero weights) and must hold
  // zeros, and the array keeps whatever the previous call left: clear the
  // activation regions first (a fill kernel, no allocation).
  const size_t words = (kNoAlloc.bytes - kNoAlloc.act_begin) / 4;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(nblk(words, 256)), dim3(256), 0, s,
                     reinterpret_cast<uint32_t *>(base + kNoAlloc.act_begin), words);
  Net n;
  n.D = D;
  n.H = H;
  n.W = W;
  int64_t off = 0;
  for (int l = 0; l < kLayers; ++l)
  {
    const LayerSpec & sp = kSpec[l];
    PackedLayer & L = n.layer[l];
    const int cin = sp.cin_a + sp.cin_b;
    L.cout = sp.cout;
    L.relu = sp.relu;
    L.taps = sp.taps;
    if (sp.taps == 1)
    {
      // Head: its W[0][0][ci] and bias are contiguous in the flat parameters.
      L.bias = const_cast<float *>(d_params + off);
      L.cin_a_p = cin;
    }
    else
    {
      L.cin_a_p = pad8(sp.cin_a);
      L.cin_b_p = sp.cin_b ? pad8(sp.cin_b) : 0;
      L.cout16 = (sp.cout + 15) / 16 * 16;
      L.kpad = (int)layer_kpad(sp);
      L.w = reinterpret_cast<uint16_t *>(base + kNoAlloc.w[l]);
      L.bias = const_cast<float *>(d_params + off + (int64_t)sp.cout * 27 * cin);
      const int nw = L.cout16 * L.kpad;
      hipLaunchKernelGGL(pack_layer_kernel, dim3(nblk(nw, 256)), dim3(256), 0, s, d_params + off,
                         sp.cout, sp.cin_a, sp.cin_b, L.cin_a_p, L.cin_a_p + L.cin_b_p, L.kpad,
                         nw, L.w);
    }
    off += layer_params(sp);
  }
  auto act = [base](size_t o) { return reinterpret_cast<uint16_t *>(base + o); };
  n.feat = act(kNoAlloc.feat);
  n.e1a = act(kNoAlloc.e1a);
  n.e1b = act(kNoAlloc.e1b);
  n.p1 = act(kNoAlloc.p1);
  n.e2a = act(kNoAlloc.e2a);
  n.e2b = act(kNoAlloc.e2b);
  n.p2 = act(kNoAlloc.p2);
  n.ba = act(kNoAlloc.ba);
  n.bb = act(kNoAlloc.bb);
  n.d2a = act(kNoAlloc.d2a);
  n.d2b = act(kNoAlloc.d2b);
  n.d1a = act(kNoAlloc.d1a);
  n.d1b = act(kNoAlloc.d1b);
  n.ready = true;
  hipLaunchKernelGGL(features_f32_kernel, dim3(nblk(V, 256)), dim3(256), 0, s, d_features,
                     (int)V, n.feat);
  if ((rc = body(n, s)))
    return rc;
  hipLaunchKernelGGL(head_raw_kernel, dim3(nblk(V, 256)), dim3(256), 0, s, n.d1b, pad8(C1), C1,
                     n.layer[10].bias, (int)V, d_out);
  DOMA_HIP(hipGetLastError());
  return DOMA_OK;
}

} // namespace net
} // namespace doma

extern "C" int
doma_net_forward_noalloc(const float * d_params, int64_t count, int32_t D, int32_t H, int32_t W,
                         const float * d_features, float * d_out, void * stream)
{
  using namespace doma;
  if (!d_params || !d_features || !d_out || count != net::total_params())
    return fail(DOMA_ERR_INVALID, "doma_net_forward_noalloc: bad arguments");
  return net::net_forward_noalloc(d_params, D, H, W, d_features, d_out,
                                  static_cast<hipStream_t>(stream));
}
