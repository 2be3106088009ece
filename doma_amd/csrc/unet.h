// unet.h -- the learned source-iteration correction: a 3-level 3D U-Net on
// BF16 MFMA (FP32 accumulate) that maps the last two SI increments of a group
// (the SI residual history) plus the group's sigma_t and in-group scattering
// ratio to a refinement of the geometric-series extrapolation of the remaining
// error phi_converged - phi.  Host-side
// description shared by the C ABI, the kernels (unet.hip) and the trainer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace doma
{
namespace net
{
// Channel widths of the three levels and the layer list.  Layers 0-9 are 3x3x3
// convolutions (zero padding 1) followed by ReLU; layer 10 is the 1x1x1 head.
// Input channels, per cell: d_k / s, d_(k-1) / s (s = rms of d_k), log sigma_t,
// log 1/(1 - c_gg + 1e-4).
constexpr int C1 = 12, C2 = 24, C3 = 32, CIN = 4;
constexpr int kLayers = 11;

struct LayerSpec
{
  int cin_a;  // channels read at full resolution (skip or plain input)
  int cin_b;  // channels read from the coarser level, upsampled x2 (0: none)
  int cout;
  int taps;   // 27 or 1
  int relu;
};

// enc1a enc1b | pool | enc2a enc2b | pool | bot_a bot_b | dec2a dec2b | dec1a dec1b | head
constexpr LayerSpec kSpec[kLayers] = {
    {CIN, 0, C1, 27, 1}, {C1, 0, C1, 27, 1}, {C1, 0, C2, 27, 1}, {C2, 0, C2, 27, 1},
    {C2, 0, C3, 27, 1},  {C3, 0, C3, 27, 1}, {C2, C3, C2, 27, 1}, {C2, 0, C2, 27, 1},
    {C1, C2, C1, 27, 1}, {C1, 0, C1, 27, 1}, {C1, 0, 1, 1, 0},
};

// Flat FP32 parameter vector: per layer W[cout][tap][cin_a + cin_b] then bias[cout].
constexpr int64_t layer_params(const LayerSpec & s)
{
  return (int64_t)s.cout * s.taps * (s.cin_a + s.cin_b) + s.cout;
}
constexpr int64_t total_params()
{
  int64_t n = 0;
  for (const auto & s : kSpec)
    n += layer_params(s);
  return n;
}

constexpr int pad8(int c) { return (c + 7) / 8 * 8; }

// Device-side packed layer: W as BF16 [cout16][kpad], K index = tap * cin_p + ci
// where cin_p = pad8(cin_a) + pad8(cin_b) (padding slots carry zero weights).
struct PackedLayer
{
  uint16_t * w = nullptr;
  float * bias = nullptr;
  int cin_a_p = 0, cin_b_p = 0, cout = 0, cout16 = 0, kpad = 0, taps = 27, relu = 1;
};

// Activation buffers of one network instance (BF16, channels-last, channel
// stride padded to 8).
struct Net
{
  bool ready = false;
  int every = 0;  // apply every `every` source iterations (0: off)
  int D = 0, H = 0, W = 0;
  PackedLayer layer[kLayers];
  uint16_t * feat = nullptr;  // [V][8]
  uint16_t * e1a = nullptr, * e1b = nullptr;  // [V][16]
  uint16_t * p1 = nullptr;                    // [V/8][16]
  uint16_t * e2a = nullptr, * e2b = nullptr;  // [V/8][24]
  uint16_t * p2 = nullptr;                    // [V/64][24]
  uint16_t * ba = nullptr, * bb = nullptr;    // [V/64][32]
  uint16_t * d2a = nullptr, * d2b = nullptr;  // [V/8][24]
  uint16_t * d1a = nullptr, * d1b = nullptr;  // [V][16]
  double * red = nullptr;                     // reduction partials
  double * scale = nullptr;                   // [2] rms of d_k, extrapolation gain g
};

// Packs the flat FP32 parameters into per-layer BF16 device layouts and sizes
// the activation buffers for a D x H x W grid (each divisible by 4).
int net_setup(Net & n, const float * params, int D, int H, int W);
void net_free(Net & n);

// corr[c] = gain * (dlast[c] + s * U-Net(dlast / s, dprev / s, log sigma_t,
// log 1/(1 - c_gg + 1e-4))) for group g, with s = rms(dlast), rho = rms(dlast) /
// rms(dprev) clamped to [0, 0.999] and gain = rho / (1 - rho): the network
// refines the geometric-series extrapolation gain * dlast of the remaining error
// (a zero network leaves the bare extrapolation).
// dlast and dprev are the latest two increments of plain source iterations;
// mat/sigt/sigs as in IterArgs.
int net_apply(Net & n, const double * dlast, const double * dprev, const int32_t * mat,
              const double * sigt, const double * sigs, int G, int g, double * corr,
              hipStream_t s);

// Forward pass on given features (testing): feat FP32 [V][CIN] -> raw network
// output FP32 [V] (no scale, no gain).
int net_forward_features(Net & n, const float * feat_dev, float * out_dev, hipStream_t s);

// Hand-written training on batches of `batch` cubic crops of edge `crop`
// (stacked along z): BF16 MFMA forward and data gradients, FP32 weight
// gradients, relative-L2 loss sum (p - y)^2 / sum y^2, global-norm clipping at 1
// and Adam (0.9, 0.999, 1e-8) on FP32 master weights.
struct Trainer;
int trainer_create(int crop, int batch, const float * params, Trainer ** out);
void trainer_free(Trainer * t);
// feats [batch][crop^3][CIN], targets [batch][crop^3] on the host.  update =
// false computes the loss and the gradient only (grad_out may be NULL).
int trainer_run(Trainer * t, const float * feats, const float * targets, float lr, bool update,
                float * loss, float * grad_out);
int trainer_params(Trainer * t, float * params);
} // namespace net
} // namespace doma
