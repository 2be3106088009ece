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

#include <cmath>
#include <cstring>

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
  int D, H, W, V;
  const uint16_t * w;  // [cout16][kpad]
  const float * bias;  // [cout]
  int cout, kpad, cin_p;
  uint16_t * out;      // [V][out_stride]
  int out_stride;
  int relu;
};

template <int NT, int MT>
__global__ __launch_bounds__(RT) void
conv3_kernel(ConvArgs p)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const int base = (blockIdx.x * (RT / 64) + wave) * 16 * MT;
  const int H2 = p.H >> 1, W2 = p.W >> 1;
  int vz[MT], vy[MT], vx[MT];
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
      const int z = vz[mt] + dz, y = vy[mt] + dy, x = vx[mt] + dx;
      if (tap < 27 && vin[mt] && z >= 0 && z < p.D && y >= 0 && y < p.H && x >= 0 && x < p.W)
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

// 2x2x2 average pooling, channels-last.
__global__ void
pool_kernel(const uint16_t * in, uint16_t * out, int D, int H, int W, int C)
{
  const int D2 = D >> 1, H2 = H >> 1, W2 = W >> 1;
  const size_t n = (size_t)D2 * H2 * W2 * C;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int c = (int)(i % C);
  const size_t v = i / C;
  const int x = (int)(v % W2), y = (int)((v / W2) % H2), z = (int)(v / ((size_t)W2 * H2));
  float s = 0.f;
  for (int d = 0; d < 8; ++d)
  {
    const int zz = 2 * z + (d >> 2), yy = 2 * y + ((d >> 1) & 1), xx = 2 * x + (d & 1);
    s += bf2f(in[((size_t)(zz * H + yy) * W + xx) * C + c]);
  }
  out[i] = f2bf(0.125f * s);
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

__global__ __launch_bounds__(RT) void
rms_final_kernel(const double * red, int nblocks, int N, double * scale)
{
  __shared__ double sh[RT];
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += RT)
    s += red[b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = RT / 2; k > 0; k >>= 1)
  {
    if ((int)threadIdx.x < k)
      sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
  {
    const double r = sqrt(sh[0] / N);
    scale[0] = r > 0.0 ? r : 1.0;
  }
}

// Network input of group g: the increment scaled by its rms, log sigma_t and
// log 1/(1 - c_gg + 1e-4) with c_gg = sigma_s,gg / sigma_t.
__global__ void
features_kernel(const double * d, const double * scale, const int32_t * mat, const double * sigt,
                const double * sigs, int G, int g, int N, uint16_t * feat)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N)
    return;
  const int m = mat[c];
  const double st = sigt[(size_t)m * G + g];
  const double sgg = sigs[((size_t)m * G + g) * G + g];
  const double cgg = st > 0.0 ? sgg / st : 0.0;
  const float f0 = (float)(d[c] / scale[0]);
  const float f1 = (float)log(st > 1e-6 ? st : 1e-6);
  const float f2 = (float)log(1.0 / (1.0 - cgg + 1e-4));
  uint16_t * o = feat + (size_t)c * 8;
  o[0] = f2bf(f0);
  o[1] = f2bf(f1);
  o[2] = f2bf(f2);
  for (int k = 3; k < 8; ++k)
    o[k] = 0;
}

__global__ void
features_f32_kernel(const float * in, int N, uint16_t * feat)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N)
    return;
  for (int k = 0; k < 8; ++k)
    feat[(size_t)c * 8 + k] = k < 3 ? f2bf(in[(size_t)c * 3 + k]) : 0;
}

// 1x1x1 head: y = w . h + b.  Raw output for the stand-alone forward pass.
__device__ inline float
head_value(const uint16_t * h, int stride, int C, const float * wb, int c)
{
  float s = wb[C];
  for (int k = 0; k < C; ++k)
    s += wb[k] * bf2f(h[(size_t)c * stride + k]);
  return s;
}

__global__ void
head_raw_kernel(const uint16_t * h, int stride, int C, const float * wb, int N, float * out)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < N)
    out[c] = head_value(h, stride, C, wb, c);
}

// Correction: e = scale * y / (1 - c_gg + 1e-4).  The network predicts the error
// in units of the infinite-medium gain, which keeps its output O(1) from weakly
// to strongly scattering cells.
__global__ void
head_corr_kernel(const uint16_t * h, int stride, int C, const float * wb, const double * scale,
                 const int32_t * mat, const double * sigt, const double * sigs, int G, int g,
                 int N, double * out)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N)
    return;
  const int m = mat[c];
  const double st = sigt[(size_t)m * G + g];
  const double cgg = st > 0.0 ? sigs[((size_t)m * G + g) * G + g] / st : 0.0;
  out[c] = scale[0] * (double)head_value(h, stride, C, wb, c) / (1.0 - cgg + 1e-4);
}

inline unsigned
nblk(size_t n, int t)
{
  return (unsigned)((n + t - 1) / t);
}

int
conv(const PackedLayer & L, const uint16_t * a, const uint16_t * b, int D, int H, int W,
     uint16_t * out, int out_stride, hipStream_t s)
{
  ConvArgs p{};
  p.a = a;
  p.ca = L.cin_a_p;
  p.b = b;
  p.cb = L.cin_b_p;
  p.D = D;
  p.H = H;
  p.W = W;
  p.V = D * H * W;
  p.w = L.w;
  p.bias = L.bias;
  p.cout = L.cout;
  p.kpad = L.kpad;
  p.cin_p = L.cin_a_p + L.cin_b_p;
  p.out = out;
  p.out_stride = out_stride;
  p.relu = L.relu;
  constexpr int MT = 2;
  const unsigned grid = nblk(p.V, (RT / 64) * 16 * MT);
  if (L.cout16 == 16)
    hipLaunchKernelGGL((conv3_kernel<1, MT>), dim3(grid), dim3(RT), 0, s, p);
  else if (L.cout16 == 32)
    hipLaunchKernelGGL((conv3_kernel<2, MT>), dim3(grid), dim3(RT), 0, s, p);
  else
    return fail(DOMA_ERR_INVALID, "unsupported layer width");
  return DOMA_OK;
}

int
pool(const uint16_t * in, uint16_t * out, int D, int H, int W, int C, hipStream_t s)
{
  const size_t n = (size_t)(D / 2) * (H / 2) * (W / 2) * C;
  hipLaunchKernelGGL(pool_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, in, out, D, H, W, C);
  return DOMA_OK;
}

// Layers 0-9 on the features in n.feat; leaves the last hidden layer in n.d1b.
int
body(Net & n, hipStream_t s)
{
  const int D = n.D, H = n.H, W = n.W;
  const int D2 = D / 2, H2 = H / 2, W2 = W / 2, D4 = D / 4, H4 = H / 4, W4 = W / 4;
  const int s1 = pad8(C1), s2 = pad8(C2), s3 = pad8(C3);
  int rc;
  if ((rc = conv(n.layer[0], n.feat, nullptr, D, H, W, n.e1a, s1, s)) ||
      (rc = conv(n.layer[1], n.e1a, nullptr, D, H, W, n.e1b, s1, s)) ||
      (rc = pool(n.e1b, n.p1, D, H, W, s1, s)) ||
      (rc = conv(n.layer[2], n.p1, nullptr, D2, H2, W2, n.e2a, s2, s)) ||
      (rc = conv(n.layer[3], n.e2a, nullptr, D2, H2, W2, n.e2b, s2, s)) ||
      (rc = pool(n.e2b, n.p2, D2, H2, W2, s2, s)) ||
      (rc = conv(n.layer[4], n.p2, nullptr, D4, H4, W4, n.ba, s3, s)) ||
      (rc = conv(n.layer[5], n.ba, nullptr, D4, H4, W4, n.bb, s3, s)) ||
      (rc = conv(n.layer[6], n.e2b, n.bb, D2, H2, W2, n.d2a, s2, s)) ||
      (rc = conv(n.layer[7], n.d2a, nullptr, D2, H2, W2, n.d2b, s2, s)) ||
      (rc = conv(n.layer[8], n.e1b, n.d2b, D, H, W, n.d1a, s1, s)) ||
      (rc = conv(n.layer[9], n.d1a, nullptr, D, H, W, n.d1b, s1, s)))
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
      (rc = dalloc0(&n.scale, 1)))
    return rc;
  n.ready = true;
  return DOMA_OK;
}

int
net_apply(Net & n, const double * dlast, const int32_t * mat, const double * sigt,
          const double * sigs, int G, int g, double * corr, hipStream_t s)
{
  if (!n.ready)
    return fail(DOMA_ERR_STATE, "the learned correction has no weights");
  const int N = n.D * n.H * n.W;
  const int nb = 1024;
  hipLaunchKernelGGL(sumsq_kernel, dim3(nb), dim3(RT), 0, s, dlast, N, n.red);
  hipLaunchKernelGGL(rms_final_kernel, dim3(1), dim3(RT), 0, s, n.red, nb, N, n.scale);
  hipLaunchKernelGGL(features_kernel, dim3(nblk(N, 256)), dim3(256), 0, s, dlast, n.scale, mat,
                     sigt, sigs, G, g, N, n.feat);
  int rc = body(n, s);
  if (rc)
    return rc;
  hipLaunchKernelGGL(head_corr_kernel, dim3(nblk(N, 256)), dim3(256), 0, s, n.d1b, pad8(C1), C1,
                     n.layer[10].bias, n.scale, mat, sigt, sigs, G, g, N, corr);
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
} // namespace net
} // namespace doma
