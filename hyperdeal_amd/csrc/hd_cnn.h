// hd_cnn.h -- the convolutional network behind the learned initial guess of
// the Vlasov-Poisson potential solve (host-side declarations).
//
// Network: four 3x3 convolutions with periodic padding on an H x W image,
// channels 1 -> 32 -> 32 -> 32 -> 1, leaky ReLU (slope 1/8) after the first
// three.  FP32 master parameters and Adam moments; the convolutions read
// BF16 weights and accumulate in FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hd
{
namespace cnn
{
constexpr int kChannels = 32;
constexpr int kLayers = 4;
constexpr int kTile = 16; // image edges must be multiples of this
} // namespace cnn

struct CnnNet
{
  int H = 0, W = 0;
  float *p = nullptr;     // FP32 master parameters (canonical layout)
  float *m = nullptr;     // Adam first moments
  float *v = nullptr;     // Adam second moments
  float *grad = nullptr;  // gradient of the last loss_and_grad
  uint32_t *pk = nullptr; // BF16 weights, packed per layer (see pack_offset)
  double *loss_dev = nullptr;
  float *act = nullptr; // activations and gradient buffers for act_batch
  float *part = nullptr; // weight-gradient partial sums per block
  int act_batch = 0;
  int steps = 0; // Adam steps taken

  // canonical layout per layer l: w[Co][Ci][3][3] then b[Co]
  static int in_channels(int l);
  static int out_channels(int l);
  static int param_offset_w(int l);
  static int param_offset_b(int l);
  static int n_params();
  // packed BF16 weights per layer: the forward copy [Ci][9][Co'] and the
  // copy for the input gradient [Co][9][Ci'] with the stencil reversed, Co'
  // and Ci' rounded up to even so that one 32-bit word holds two channels
  static int pack_offset_fwd(int l);
  static int pack_offset_bwd(int l);
  static int n_packed();
  static bool shape_ok(int h, int w);

  hipError_t init(int h, int w, uint64_t seed);
  void release();
  // master parameters -> packed BF16 copies
  hipError_t pack(hipStream_t s);
  hipError_t reserve(int batch);
  // out = net(in); in, out: batch x H x W
  hipError_t forward(const float *in, float *out, int batch, hipStream_t s);
  // loss = mean (net(in) - target)^2 into loss_dev, gradient into grad
  hipError_t loss_and_grad(const float *in, const float *target, int batch,
                           hipStream_t s);
  // one Adam step (beta 0.9, 0.999, eps 1e-8) with the held gradient
  hipError_t adam(float lr, hipStream_t s);
};

// x-space nodal values (2D, cell-major, x0 cells fastest, N x N nodes per
// cell, node index i0 + N i1) <-> an (n1 N) x (n0 N) image, row = x1
//   v = u - sub (sub may be null), image = (v - mean v) / scale,
//   scale = rms(v - mean v) computed on the device into *scale (0 when v
//   is constant, then the image is 0)
hipError_t cnn_dofs_to_image(const double *u, const double *sub, int n0,
                             int n1, int N, float *image, double *scale,
                             hipStream_t s);
//   u = scale * image, scale read from the device, continuous across cell
//   faces (nodes on a face or corner get the mean of their 2 or 4 pixels)
hipError_t cnn_image_to_dofs(const float *image, const double *scale, int n0,
                             int n1, int N, double *u, hipStream_t s);
//   a training batch: dst[j] = src[i_j], dst2[j] = src2[i_j] for n images of
//   `plane` floats, i_j uniform over the `count` images [0, count] without
//   `skip` (skip < 0: [0, count)) from SplitMix64 (seed, counter0 + j),
//   drawn on the device
hipError_t cnn_sample(const float *src, const float *src2, uint64_t seed,
                      uint64_t counter0, int count, int skip, int n,
                      long long plane, float *dst, float *dst2,
                      hipStream_t s);

uint16_t cnn_to_bf16(float x);
} // namespace hd
