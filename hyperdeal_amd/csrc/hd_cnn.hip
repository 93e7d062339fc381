// hd_cnn.hip -- a small convolutional network that maps the x-space charge
// density of the Vlasov-Poisson problem to an initial guess for the
// potential (no counterpart in the reference, which starts CG from the
// previous potential, vlasov_poisson/include/application.h:575-576).
//
// Four 3x3 convolutions with periodic padding on an H x W image of nodal
// values, channels 1 -> 32 -> 32 -> 32 -> 1, leaky ReLU (slope 1/8) after
// the first three.  Weights are read in BF16, two channels per 32-bit word,
// through scalar loads (they are uniform over a workgroup), and every
// product is accumulated in FP32; biases, activations and the Adam state are
// FP32.  Convolutions run as direct stencils over 16 x 16 pixel tiles staged
// in LDS; the weight gradient is one workgroup per (output, input) channel
// pair reducing over the pixels in a fixed order, so every result is
// deterministic.
#include "hd_cnn.h"
#include "hd_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace hd
{
namespace cnn
{
constexpr int kHalo = kTile + 2;
constexpr float kSlope = 0.125f;

template <class T>
using cptr = const __attribute__((address_space(4))) T *;

// FP32 -> BF16, round to nearest even (NaN stays NaN)
__host__ __device__ inline uint16_t to_bf16(float x)
{
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u)
    return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bits_to_float(uint32_t u)
{
  float x;
  memcpy(&x, &u, 4);
  return x;
}

__device__ inline int wrap(int i, int n)
{
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

// out[b][o] = act(bias[o] + sum_{i,k} w[i][k][o] in[b][i] shifted by k),
// w packed as [CIN][9][COUT rounded up to even] BF16.  Grid
// (W / 16, H / 16, batch * COUT / OT), one output pixel and OT channels per
// thread.
template <int CIN, int COUT, bool ACT>
__global__ __launch_bounds__(kTile *kTile) void conv_kernel(
  const float *__restrict__ in, const uint32_t *__restrict__ w,
  const float *__restrict__ bias, float *__restrict__ out, int H, int W)
{
  constexpr int OT = COUT < 8 ? COUT : 8;
  constexpr int NG = COUT / OT;
  constexpr int WPR = (COUT + (COUT & 1)) / 2; // words per (i, k)
  constexpr int IT = CIN < 8 ? CIN : 8;
  static_assert(COUT % OT == 0 && CIN % IT == 0, "channel tiling");
  __shared__ float tile[IT][kHalo][kHalo];
  const cptr<uint32_t> wc = (cptr<uint32_t>)w;
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int b = blockIdx.z / NG, og = blockIdx.z % NG;
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  const long long plane = (long long)H * W;
  float acc[OT];
#pragma unroll
  for (int o = 0; o < OT; ++o)
    acc[o] = 0.0f;
  for (int i0 = 0; i0 < CIN; i0 += IT)
    {
      for (int e = threadIdx.x; e < IT * kHalo * kHalo; e += kTile * kTile)
        {
          const int i = e / (kHalo * kHalo), r = e % (kHalo * kHalo);
          const int yy = wrap(y0 + r / kHalo - 1, H);
          const int xx = wrap(x0 + r % kHalo - 1, W);
          tile[i][r / kHalo][r % kHalo] =
            in[((long long)b * CIN + i0 + i) * plane + (long long)yy * W + xx];
        }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int k = 0; k < 9; ++k)
          {
            const float v = tile[i][ty + k / 3][tx + k % 3];
            const int base = ((i0 + i) * 9 + k) * WPR + og * (OT / 2);
            if constexpr (OT == 1)
              acc[0] += bits_to_float(wc[base] << 16) * v;
            else
#pragma unroll
              for (int j = 0; j < OT / 2; ++j)
                {
                  const uint32_t u = wc[base + j];
                  acc[2 * j] += bits_to_float(u << 16) * v;
                  acc[2 * j + 1] += bits_to_float(u & 0xffff0000u) * v;
                }
          }
      __syncthreads();
    }
  const int x = x0 + tx, y = y0 + ty;
#pragma unroll
  for (int o = 0; o < OT; ++o)
    {
      const int oc = og * OT + o;
      float v = acc[o] + (bias ? ((cptr<float>)bias)[oc] : 0.0f);
      if (ACT)
        v = v > 0.0f ? v : kSlope * v;
      out[((long long)b * COUT + oc) * plane + (long long)y * W + x] = v;
    }
}

// ---------------------------------------------------------------------------
// The 32 -> 32 layers on the matrix cores.  A convolution is a GEMM over
// K = 32 input channels x 9 taps: out[o][pixel] = sum_k W[o][k] in[k][pixel]
// (in shifted by the tap).  v_mfma_f32_16x16x4_f32 takes FP32 operands, so
// the BF16 weights (exact in FP32) and the FP32 activations enter unrounded
// and every product is accumulated in FP32, as in the direct stencils.
// Operand layout of one 16x16x4 step (64 lanes, l = lane):
//   A[m][k]: lane m + 16 k supplies A[m][k]   (m < 16, k < 4)
//   B[k][n]: lane n + 16 k supplies B[k][n]
//   D[m][n]: lane n + 16 (m / 4) holds D[m][n] in element m % 4
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

__device__ inline mfma_f4 mfma16x16x4(float a, float b, mfma_f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int kMTX = 16, kMTY = 8; // pixels per block (x, y)
// LDS strides chosen so that the four 16-lane groups of a read (one channel
// each) fall on disjoint banks (stride = 16 or 48 mod 64 floats)
constexpr int kTileStride = 208; // >= (kMTY + 2) * (kMTX + 2) = 180
constexpr int kWStride = 304;    // >= 9 * 32 = 288

// out = act(bias + conv(in)) for 32 -> 32 channels; grid (W / 16, H / 8,
// batch), four waves per block, wave w owns rows 2w, 2w + 1 of the block's
// 16 x 8 pixels and all 32 output channels (2 x 2 tiles of 16 x 16).  The
// block stages all 32 input channels (with a one-pixel periodic halo) and
// all weights (as FP32) in LDS once, then runs its 9 x 8 x 4 MFMA steps
// without further barriers (dynamic LDS: kConvLds bytes).
constexpr int kConvLds =
  kChannels * (kTileStride + kWStride) * (int)sizeof(float);
template <bool ACT, int OH>
__global__ __launch_bounds__(256) void conv32_mfma_kernel(
  const float *__restrict__ in, const uint32_t *__restrict__ w,
  const float *__restrict__ bias, float *__restrict__ out, int H, int W)
{
  constexpr int C = kChannels;
  constexpr int WPR = C / 2;        // packed words per (input channel, tap)
  constexpr int WPH = WPR / OH;     // of them for this block's channels
  constexpr int TO = 2 / OH;        // 16-channel output tiles per block
  constexpr int HW = (kMTY + 2) * (kMTX + 2);
  extern __shared__ float sm[];
  float *tile = sm;                 // [C][kTileStride]
  float *wl = sm + C * kTileStride; // [C][kWStride]: wl[i][tap][o]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n16 = lane & 15, k4 = lane >> 4;
  const int b = blockIdx.z / OH, half = blockIdx.z % OH;
  const int x0 = blockIdx.x * kMTX, y0 = blockIdx.y * kMTY;
  const long long plane = (long long)H * W;
  for (int e = threadIdx.x; e < C * HW; e += 256)
    {
      const int i = e / HW, r = e % HW;
      const int yy = wrap(y0 + r / (kMTX + 2) - 1, H);
      const int xx = wrap(x0 + r % (kMTX + 2) - 1, W);
      tile[i * kTileStride + r] =
        in[((long long)b * C + i) * plane + (long long)yy * W + xx];
    }
  for (int e = threadIdx.x; e < C * 9 * WPH; e += 256)
    {
      const int i = e / (9 * WPH), r = e % (9 * WPH);
      const int tap = r / WPH, pr = half * WPH + r % WPH;
      const uint32_t u = w[(i * 9 + tap) * WPR + pr];
      wl[i * kWStride + tap * C + 2 * pr] = bits_to_float(u << 16);
      wl[i * kWStride + tap * C + 2 * pr + 1] = bits_to_float(u & 0xffff0000u);
    }
  __syncthreads();
  mfma_f4 acc[2][TO];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int t = 0; t < TO; ++t)
      acc[r][t] = mfma_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
    {
      const int dy = tap / 3, dx = tap % 3;
#pragma unroll 2
      for (int q = 0; q < C / 4; ++q)
        {
          const int ci = q * 4 + k4;
          // A = weights (m = output channel), B = pixels (n = x)
          float a[TO];
#pragma unroll
          for (int t = 0; t < TO; ++t)
            a[t] = wl[ci * kWStride + tap * C + 16 * (half + t) + n16];
#pragma unroll
          for (int r = 0; r < 2; ++r)
            {
              const int row = wave * 2 + r;
              const float bv =
                tile[ci * kTileStride + (row + dy) * (kMTX + 2) + n16 + dx];
#pragma unroll
              for (int t = 0; t < TO; ++t)
                acc[r][t] = mfma16x16x4(a[t], bv, acc[r][t]);
            }
        }
    }
  // lane holds D[o = 16 (half + t) + 4 k4 + e][x = n16] of rows 2 wave + r
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int t = 0; t < TO; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        {
          const int oc = 16 * (half + t) + 4 * k4 + e;
          float v = acc[r][t][e] + (bias ? ((cptr<float>)bias)[oc] : 0.0f);
          if (ACT)
            v = v > 0.0f ? v : kSlope * v;
          const int y = y0 + wave * 2 + r, x = x0 + n16;
          out[((long long)b * C + oc) * plane + (long long)y * W + x] = v;
        }
}

// out[b] = bias + sum_{i,k} w[i][k] in[b][i] shifted by k for CIN -> 1
// channels (the last layer, no activation).  A block of 256 threads owns a
// 16 x 4 pixel tile; wave g sums the channels 8 g .. 8 g + 7 (taps inner)
// for the lane's pixel with the inputs read straight from global memory
// (L2: every value is read by 9 lanes of neighbouring tiles), the weights
// through scalar loads; the four partial sums are added in wave order
// through LDS, so the result is deterministic
template <int CIN>
__global__ __launch_bounds__(256) void conv_to1_kernel(
  const float *__restrict__ in, const uint32_t *__restrict__ w,
  const float *__restrict__ bias, float *__restrict__ out, int H, int W)
{
  constexpr int G = 4, CG = CIN / G;
  static_assert(CIN % G == 0, "channel groups");
  __shared__ float part[G][64];
  const cptr<uint32_t> wc = (cptr<uint32_t>)w;
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int tiles_x = W / 16, tiles = tiles_x * (H / 4);
  const int b = blockIdx.x / tiles, tl = blockIdx.x % tiles;
  const int y = (tl / tiles_x) * 4 + lane / 16;
  const int x = (tl % tiles_x) * 16 + lane % 16;
  const int plane = H * W;
  const int rows[3] = {wrap(y - 1, H) * W, y * W, wrap(y + 1, H) * W};
  const int cols[3] = {wrap(x - 1, W), x, wrap(x + 1, W)};
  const float *ib = in + ((long long)b * CIN + g * CG) * plane;
  float v[CG][9];
#pragma unroll
  for (int i = 0; i < CG; ++i)
#pragma unroll
    for (int k = 0; k < 9; ++k)
      v[i][k] = ib[(long long)i * plane + rows[k / 3] + cols[k % 3]];
  float acc = 0.0f;
#pragma unroll
  for (int i = 0; i < CG; ++i)
#pragma unroll
    for (int k = 0; k < 9; ++k)
      acc += bits_to_float(wc[(g * CG + i) * 9 + k] << 16) * v[i][k];
  part[g][lane] = acc;
  __syncthreads();
  if (g == 0)
    {
      float t = part[0][lane];
#pragma unroll
      for (int j = 1; j < G; ++j)
        t += part[j][lane];
      out[(long long)b * plane + y * W + x] =
        t + (bias ? ((cptr<float>)bias)[0] : 0.0f);
    }
}

// Weight gradient of a 32 -> 32 layer: gw[o][i][tap] = sum over the batch
// and pixels of g[o][pixel] in[i][pixel + tap]; a GEMM with M = 32 (o),
// N = 288 (i, tap), K = pixels.  Block p sums the 16 x 16 pixel tiles p,
// p + gridDim.x, ... into part[p][o][i][tap] (wave w owns the N tiles
// w, w + 4, ... of the 18, for both M tiles); the blocks' partials are then
// added in a fixed order (part_reduce_kernel), so the result is
// deterministic.
constexpr int kGStride = 272;   // 256 pixels of one gradient channel, +16
constexpr int kWgradOut = kChannels * kChannels * 9 + kChannels; // per block
constexpr int kInStride = 336;  // 18 x 18 = 324 input pixels, 336 = 16 mod 64
__global__ __launch_bounds__(256) void wgrad32_mfma_kernel(
  const float *__restrict__ in, const float *__restrict__ g,
  float *__restrict__ part, int B, int H, int W)
{
  constexpr int C = kChannels;
  constexpr int NT = C * 9 / 16; // 18 N tiles
  extern __shared__ float sm[];
  float *gs = sm;                  // [C][kGStride]
  float *is = sm + C * kGStride;   // [C][kInStride]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n16 = lane & 15, k4 = lane >> 4;
  const int tx = W / 16, ty = H / 16;
  const long long n_tiles = (long long)B * tx * ty;
  const long long plane = (long long)H * W;
  // this lane's B column in every N tile it owns: (input channel, tap)
  constexpr int NW = (NT + 3) / 4; // N tiles per wave (at most)
  mfma_f4 acc[2][NW];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int j = 0; j < NW; ++j)
      acc[m][j] = mfma_f4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  for (long long t = blockIdx.x; t < n_tiles; t += gridDim.x)
    {
      const int b = (int)(t / (tx * ty));
      const int rem = (int)(t % (tx * ty));
      const int y0 = (rem / tx) * 16, x0 = (rem % tx) * 16;
      __syncthreads(); // the previous tile's LDS reads are done
      for (int e = threadIdx.x; e < C * 256; e += 256)
        {
          const int o = e / 256, p = e % 256;
          gs[o * kGStride + p] = g[((long long)b * C + o) * plane +
                                   (long long)(y0 + p / 16) * W + x0 + p % 16];
        }
      for (int e = threadIdx.x; e < C * 324; e += 256)
        {
          const int i = e / 324, r = e % 324;
          const int yy = wrap(y0 + r / 18 - 1, H), xx = wrap(x0 + r % 18 - 1, W);
          is[i * kInStride + r] =
            in[((long long)b * C + i) * plane + (long long)yy * W + xx];
        }
      __syncthreads();
      // this thread's share of the bias gradient: channel t / 8, pixels
      // 32 (t % 8) .. + 31 of the tile
      {
        const float *gr = gs + (threadIdx.x / 8) * kGStride + 32 * (threadIdx.x % 8);
#pragma unroll 8
        for (int j = 0; j < 32; ++j)
          bsum += gr[j];
      }
      // K steps of 4 pixels: pixel p = 4 s + k4 of the tile
      for (int s = 0; s < 64; ++s)
        {
          const int p = 4 * s + k4;
          const int py = p / 16, px = p % 16;
          // A[m = o][k = pixel] = g[o][pixel]
          const float a0 = gs[n16 * kGStride + p];
          const float a1 = gs[(16 + n16) * kGStride + p];
#pragma unroll
          for (int j = 0; j < NW; ++j)
            {
              const int nt = wave + 4 * j;
              if (nt < NT)
                {
                  // B[k = pixel][n = (i, tap)] = in[i][pixel + tap]
                  const int col = nt * 16 + n16;
                  const int i = col / 9, tap = col % 9;
                  const float bv =
                    is[i * kInStride + (py + tap / 3) * 18 + px + tap % 3];
                  acc[0][j] = mfma16x16x4(a0, bv, acc[0][j]);
                  acc[1][j] = mfma16x16x4(a1, bv, acc[1][j]);
                }
            }
        }
    }
  // D[o = 16 m + 4 k4 + e][col = 16 nt + n16] -> part[block][o][col]; the
  // bias gradient's block partials after them: part[block][C * C * 9 + o]
  float *pb = part + (long long)blockIdx.x * kWgradOut;
  {
    __shared__ float bred[256];
    bred[threadIdx.x] = bsum;
    __syncthreads();
    if (threadIdx.x < C)
      {
        float t = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          t += bred[threadIdx.x * 8 + j];
        pb[C * C * 9 + threadIdx.x] = t;
      }
  }
#pragma unroll
  for (int j = 0; j < NW; ++j)
    {
      const int nt = wave + 4 * j;
      if (nt < NT)
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            pb[(16 * m + 4 * k4 + e) * (C * 9) + nt * 16 + n16] = acc[m][j][e];
    }
}

// out[e] = sum over n_part partials (stride n_out) of element e, in a fixed
// order: four threads per element each sum every fourth partial, then the
// four sums are added in order; e < n_w goes to gw[e], the rest to
// gb[e - n_w]
__global__ __launch_bounds__(256) void part_reduce_kernel(
  const float *__restrict__ part, int n_part, int n_out, int n_w,
  float *__restrict__ gw, float *__restrict__ gb)
{
  __shared__ float red[4][64];
  const int o = blockIdx.x * 64 + threadIdx.x % 64, q = threadIdx.x / 64;
  float s = 0.0f;
  if (o < n_out)
    for (int p = q; p < n_part; p += 4)
      s += part[(long long)p * n_out + o];
  red[q][threadIdx.x % 64] = s;
  __syncthreads();
  if (q == 0 && o < n_out)
    {
      const int l = threadIdx.x;
      const float t = ((red[0][l] + red[1][l]) + red[2][l]) + red[3][l];
      if (o < n_w)
        gw[o] = t;
      else
        gb[o - n_w] = t;
    }
}

// Weight and bias gradients of the 1 -> 32 and 32 -> 1 layers, split over
// the pixels: block (pair (o, i), chunk) sums the nine tap products and
// (for i == 0) the bias term over its share of the pixel groups into
// part[chunk][pair][10]; part_reduce_kernel adds the chunks in order
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void wgrad_split_kernel(
  const float *__restrict__ in, const float *__restrict__ g,
  float *__restrict__ part, int B, int H, int W)
{
  __shared__ float red[10][256];
  const int pair = blockIdx.x, chunk = blockIdx.y, n_chunks = gridDim.y;
  const int o = pair / CIN, i = pair % CIN;
  const long long plane = (long long)H * W;
  const int W4 = W / 4;
  const long long groups = (long long)B * H * W4;
  const long long e0 = groups * chunk / n_chunks;
  const long long e1 = groups * (chunk + 1) / n_chunks;
  float s[10];
#pragma unroll
  for (int k = 0; k < 10; ++k)
    s[k] = 0.0f;
  for (long long e = e0 + threadIdx.x; e < e1; e += blockDim.x)
    {
      const int b = (int)(e / ((long long)H * W4));
      const int r = (int)(e % ((long long)H * W4));
      const int y = r / W4, x = (r % W4) * 4;
      const float4 gv = *reinterpret_cast<const float4 *>(
        g + ((long long)b * COUT + o) * plane + (long long)y * W + x);
      const float *ib = in + ((long long)b * CIN + i) * plane;
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
        {
          const float *row = ib + (long long)wrap(y + dy - 1, H) * W;
          const float4 mid = *reinterpret_cast<const float4 *>(row + x);
          const float v[6] = {row[wrap(x - 1, W)], mid.x, mid.y, mid.z,
                              mid.w, row[wrap(x + 4, W)]};
#pragma unroll
          for (int dx = 0; dx < 3; ++dx)
            s[dy * 3 + dx] += gv.x * v[dx] + gv.y * v[dx + 1] +
                              gv.z * v[dx + 2] + gv.w * v[dx + 3];
        }
      s[9] += (gv.x + gv.y) + (gv.z + gv.w);
    }
#pragma unroll
  for (int k = 0; k < 10; ++k)
    red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1)
    {
      if ((int)threadIdx.x < st)
#pragma unroll
        for (int k = 0; k < 10; ++k)
          red[k][threadIdx.x] += red[k][threadIdx.x + st];
      __syncthreads();
    }
  if (threadIdx.x < 10)
    part[((long long)chunk * gridDim.x + pair) * 10 + threadIdx.x] =
      red[threadIdx.x][0];
}

// the split partials, summed over the chunks in order, to the canonical
// layout: gw[o][i][tap], gb[o] (from the pairs with i == 0)
__global__ __launch_bounds__(256) void wgrad_split_reduce_kernel(
  const float *__restrict__ part, int n_chunks, int cin, int n_pairs,
  float *__restrict__ gw, float *__restrict__ gb)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_pairs * 10)
    return;
  const int pair = e / 10, k = e % 10;
  float s = 0.0f;
  for (int c = 0; c < n_chunks; ++c)
    s += part[((long long)c * n_pairs + pair) * 10 + k];
  const int o = pair / cin, i = pair % cin;
  if (k < 9)
    gw[((long long)o * cin + i) * 9 + k] = s;
  else if (i == 0)
    gb[o] = s;
}

// loss partials: g = 2 (out - target) / n and, per block, the FP64 sum of
// (out - target)^2 over its strided share; mse_finish_kernel adds the
// blocks' sums in order
constexpr int kMseBlocks = 64;
__global__ __launch_bounds__(256) void mse_partial_kernel(
  const float *__restrict__ out, const float *__restrict__ target,
  float *__restrict__ g, long long n, double *__restrict__ partial)
{
  __shared__ double red[256];
  double s = 0.0;
  const float sc = (float)(2.0 / (double)n);
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    {
      const float d = out[i] - target[i];
      g[i] = sc * d;
      s += (double)d * d;
    }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1)
    {
      if ((int)threadIdx.x < st)
        red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
  if (threadIdx.x == 0)
    partial[blockIdx.x] = red[0];
}

__global__ void mse_finish_kernel(const double *__restrict__ partial, int nb,
                                  long long n, double *__restrict__ loss)
{
  if (threadIdx.x == 0)
    {
      double s = 0.0;
      for (int b = 0; b < nb; ++b)
        s += partial[b];
      *loss = s / (double)n;
    }
}

// g *= leaky-ReLU'(pre-activation); a is the post-activation output, which
// has the sign of the pre-activation
__global__ void act_grad_kernel(float *__restrict__ g,
                                const float *__restrict__ a, long long n)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    g[i] *= a[i] > 0.0f ? 1.0f : kSlope;
}

// Adam on the FP32 master parameters (bias-corrected step sizes c1, c2)
__global__ void adam_kernel(float *__restrict__ p, float *__restrict__ m,
                            float *__restrict__ v,
                            const float *__restrict__ grad, int n, float lr,
                            float b1, float b2, float c1, float c2, float eps)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    {
      const float gr = grad[i];
      const float mi = b1 * m[i] + (1.0f - b1) * gr;
      const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
      m[i] = mi;
      v[i] = vi;
      p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// master weights w[Co][Ci][9] of one layer -> the two packed BF16 copies
__global__ void pack_kernel(const float *__restrict__ w, int Ci, int Co,
                            uint32_t *__restrict__ fwd,
                            uint32_t *__restrict__ bwd)
{
  const int wpo = (Co + 1) / 2, wpi = (Ci + 1) / 2;
  const int n_f = Ci * 9 * wpo, n_b = Co * 9 * wpi;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n_f + n_b;
       q += gridDim.x * blockDim.x)
    {
      uint32_t lo = 0, hi = 0;
      if (q < n_f)
        {
          const int pr = q % wpo, k = (q / wpo) % 9, i = q / (wpo * 9);
          const int o0 = 2 * pr, o1 = 2 * pr + 1;
          lo = to_bf16(w[(o0 * Ci + i) * 9 + k]);
          if (o1 < Co)
            hi = to_bf16(w[(o1 * Ci + i) * 9 + k]);
          fwd[q] = lo | (hi << 16);
        }
      else
        {
          const int qb = q - n_f;
          const int pr = qb % wpi, k = (qb / wpi) % 9, o = qb / (wpi * 9);
          const int i0 = 2 * pr, i1 = 2 * pr + 1;
          lo = to_bf16(w[(o * Ci + i0) * 9 + 8 - k]);
          if (i1 < Ci)
            hi = to_bf16(w[(o * Ci + i1) * 9 + 8 - k]);
          bwd[qb] = lo | (hi << 16);
        }
    }
}

// ---- x-space nodal values <-> images ----
// (an image holds fewer than 2^31 pixels: shape_ok and the callers' sizes)
__device__ inline int pixel_of(int dof, int n0, int N)
{
  const int NN = N * N;
  const int c = dof / NN;
  const int i = dof - c * NN;
  const int cy = c / n0, cx = c - cy * n0;
  const int iy = i / N, ix = i - iy * N;
  return (cy * N + iy) * (n0 * N) + cx * N + ix;
}

__device__ inline double block_sum(double v, double *red)
{
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1)
    {
      if ((int)threadIdx.x < st)
        red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
  const double r = red[0];
  __syncthreads();
  return r;
}

// v = u - sub (sub may be null): image = (v - mean v) / rms, scale = rms
__global__ __launch_bounds__(1024) void dofs_to_image_kernel(
  const double *__restrict__ u, const double *__restrict__ sub, long long n,
  int n0, int N, float *__restrict__ image, double *__restrict__ scale)
{
  __shared__ double red[1024];
  auto val = [&](long long i) { return sub ? u[i] - sub[i] : u[i]; };
  // the first kKeep values of each thread stay in registers between the
  // passes (all of them up to 16,384 DoFs: 32^2 cells of 4^2 nodes)
  constexpr int kKeep = 16;
  double keep[kKeep];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kKeep; ++k)
    {
      const long long i = threadIdx.x + (long long)k * blockDim.x;
      keep[k] = i < n ? val(i) : 0.0;
      s += keep[k];
    }
  for (long long i = threadIdx.x + (long long)kKeep * blockDim.x; i < n;
       i += blockDim.x)
    s += val(i);
  const double mean = block_sum(s, red) / (double)n;
  s = 0.0;
#pragma unroll
  for (int k = 0; k < kKeep; ++k)
    {
      const long long i = threadIdx.x + (long long)k * blockDim.x;
      if (i < n)
        s += (keep[k] - mean) * (keep[k] - mean);
    }
  for (long long i = threadIdx.x + (long long)kKeep * blockDim.x; i < n;
       i += blockDim.x)
    s += (val(i) - mean) * (val(i) - mean);
  const double rms = sqrt(block_sum(s, red) / (double)n);
  const double inv = rms > 0.0 ? 1.0 / rms : 0.0;
#pragma unroll
  for (int k = 0; k < kKeep; ++k)
    {
      const long long i = threadIdx.x + (long long)k * blockDim.x;
      if (i < n)
        image[pixel_of((int)i, n0, N)] = (float)((keep[k] - mean) * inv);
    }
  for (long long i = threadIdx.x + (long long)kKeep * blockDim.x; i < n;
       i += blockDim.x)
    image[pixel_of((int)i, n0, N)] = (float)((val(i) - mean) * inv);
  if (threadIdx.x == 0)
    *scale = rms;
}

// u = scale * image, continuous across cell faces: the Gauss-Lobatto end
// nodes of neighbouring cells sit at one point (two pixels, four at a
// corner), and every such node gets the mean of its pixels -- a guess with
// jumps at the faces would pay for them in the SIPG penalty term
__global__ void image_to_dofs_kernel(const float *__restrict__ image,
                                     const double *__restrict__ scale,
                                     long long n, int n0, int n1, int N,
                                     double *__restrict__ u)
{
  const double sc = *scale;
  const int W = n0 * N, Hh = n1 * N;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    {
      const long long p = pixel_of((int)i, n0, N);
      const int Y = (int)(p / W), X = (int)(p % W);
      const int ix = X % N, iy = Y % N;
      int xs[2] = {X, X}, ys[2] = {Y, Y};
      int nxs = 1, nys = 1;
      if (ix == 0 && N > 1)
        xs[nxs++] = X == 0 ? W - 1 : X - 1;
      else if (ix == N - 1 && N > 1)
        xs[nxs++] = X == W - 1 ? 0 : X + 1;
      if (iy == 0 && N > 1)
        ys[nys++] = Y == 0 ? Hh - 1 : Y - 1;
      else if (iy == N - 1 && N > 1)
        ys[nys++] = Y == Hh - 1 ? 0 : Y + 1;
      double acc = 0.0;
      for (int a = 0; a < nys; ++a)
        for (int b = 0; b < nxs; ++b)
          acc += (double)image[(long long)ys[a] * W + xs[b]];
      u[i] = sc * acc / (nxs * nys);
    }
}

// dst[j] = src[idx_j] for n images of `plane` floats; idx_j drawn on the
// device from SplitMix64 (seed, counter0 + j) uniformly over `count` images:
// [0, count), or [0, count] without image `skip` when skip >= 0
__global__ void gather_kernel(const float *__restrict__ src,
                              const float *__restrict__ src2, uint64_t seed,
                              uint64_t counter0, int count, int skip, int n,
                              long long plane, float *__restrict__ dst,
                              float *__restrict__ dst2)
{
  const long long total = (long long)n * plane;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       e < total; e += (long long)gridDim.x * blockDim.x)
    {
      const long long j = e / plane, p = e - j * plane;
      const double u = 0.5 * (splitmix_uniform(seed, counter0 + j) + 1.0);
      int idx = min((int)(u * count), count - 1);
      if (skip >= 0 && idx >= skip)
        ++idx;
      dst[e] = src[(long long)idx * plane + p];
      dst2[e] = src2[(long long)idx * plane + p];
    }
}
} // namespace cnn

// ---------------------------------------------------------------------------
// CnnNet

int CnnNet::in_channels(int l) { return l == 0 ? 1 : cnn::kChannels; }
int CnnNet::out_channels(int l)
{
  return l == cnn::kLayers - 1 ? 1 : cnn::kChannels;
}

int CnnNet::param_offset_w(int l)
{
  int o = 0;
  for (int j = 0; j < l; ++j)
    o += out_channels(j) * in_channels(j) * 9 + out_channels(j);
  return o;
}

int CnnNet::param_offset_b(int l)
{
  return param_offset_w(l) + out_channels(l) * in_channels(l) * 9;
}

int CnnNet::n_params() { return param_offset_w(cnn::kLayers); }

namespace
{
// blocks of the weight-gradient kernel (each sums a strided set of 16 x 16
// pixel tiles); fixed by the problem size, so the sums are deterministic
constexpr int kWgradBlocks = 256;
// pixel chunks of the small layers' weight gradients
constexpr int kSplitChunks = 32;
int wgrad_blocks(int batch, int H, int W)
{
  const long long tiles = (long long)batch * (H / 16) * (W / 16);
  return (int)std::min<long long>(tiles, kWgradBlocks);
}

int fwd_words(int l)
{
  return CnnNet::in_channels(l) * 9 * ((CnnNet::out_channels(l) + 1) / 2);
}
int bwd_words(int l)
{
  return CnnNet::out_channels(l) * 9 * ((CnnNet::in_channels(l) + 1) / 2);
}
} // namespace

int CnnNet::pack_offset_fwd(int l)
{
  int o = 0;
  for (int j = 0; j < l; ++j)
    o += fwd_words(j) + bwd_words(j);
  return o;
}

int CnnNet::pack_offset_bwd(int l) { return pack_offset_fwd(l) + fwd_words(l); }

int CnnNet::n_packed() { return pack_offset_fwd(cnn::kLayers); }

bool CnnNet::shape_ok(int h, int w)
{
  return h >= cnn::kTile && w >= cnn::kTile && h % cnn::kTile == 0 &&
         w % cnn::kTile == 0;
}

hipError_t CnnNet::init(int h, int w, uint64_t seed)
{
  H = h;
  W = w;
  const int n = n_params();
  hipError_t e = hipMalloc(&p, n * sizeof(float));
  if (e == hipSuccess)
    e = hipMalloc(&m, n * sizeof(float));
  if (e == hipSuccess)
    e = hipMalloc(&v, n * sizeof(float));
  if (e == hipSuccess)
    e = hipMalloc(&grad, n * sizeof(float));
  if (e == hipSuccess)
    e = hipMalloc(&pk, n_packed() * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMalloc(&loss_dev, sizeof(double));
  if (e != hipSuccess)
    return e;
  // uniform weights in +-sqrt(6 / fan_in) (the last layer scaled by 1/10)
  // from SplitMix64 counters 0, 1, ... in canonical order; biases zero
  std::vector<float> hp(n, 0.0f);
  uint64_t ctr = 0;
  for (int l = 0; l < cnn::kLayers; ++l)
    {
      const double lim = std::sqrt(6.0 / (in_channels(l) * 9)) *
                         (l == cnn::kLayers - 1 ? 0.1 : 1.0);
      const int ow = param_offset_w(l);
      for (int k = 0; k < out_channels(l) * in_channels(l) * 9; ++k)
        hp[ow + k] = (float)(lim * splitmix_uniform(seed, ctr++));
    }
  e = hipMemcpy(p, hp.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemset(m, 0, n * sizeof(float));
  if (e == hipSuccess)
    e = hipMemset(v, 0, n * sizeof(float));
  if (e == hipSuccess)
    e = hipMemset(grad, 0, n * sizeof(float));
  if (e == hipSuccess)
    e = pack(nullptr);
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  steps = 0;
  return e;
}

void CnnNet::release()
{
  (void)hipFree(p);
  (void)hipFree(m);
  (void)hipFree(v);
  (void)hipFree(grad);
  (void)hipFree(pk);
  (void)hipFree(loss_dev);
  (void)hipFree(act);
  (void)hipFree(part);
  p = m = v = grad = act = part = nullptr;
  pk = nullptr;
  loss_dev = nullptr;
  act_batch = 0;
}

hipError_t CnnNet::pack(hipStream_t s)
{
  for (int l = 0; l < cnn::kLayers; ++l)
    {
      const int n = fwd_words(l) + bwd_words(l);
      hipLaunchKernelGGL(cnn::pack_kernel, dim3((n + 255) / 256), dim3(256), 0,
                         s, p + param_offset_w(l), in_channels(l),
                         out_channels(l), pk + pack_offset_fwd(l),
                         pk + pack_offset_bwd(l));
    }
  return hipGetLastError();
}

hipError_t CnnNet::reserve(int batch)
{
  if (batch <= act_batch)
    return hipSuccess;
  // the buffers may still be in use by queued kernels
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess)
    return e;
  (void)hipFree(act);
  (void)hipFree(part);
  act = nullptr;
  part = nullptr;
  act_batch = 0;
  // a1, a2, a3 (C channels), out (1), g0, g1 (C channels)
  const size_t per =
    (size_t)H * W * (3 * cnn::kChannels + 1 + 2 * cnn::kChannels);
  e = hipMalloc(&act, per * batch * sizeof(float));
  // the partial sums of the weight gradients (and of the loss)
  if (e == hipSuccess)
    e = hipMalloc(&part, std::max<size_t>(
                           (size_t)wgrad_blocks(batch, H, W) * cnn::kWgradOut,
                           (size_t)kSplitChunks * cnn::kChannels * 10) *
                           sizeof(float));
  if (e == hipSuccess)
    act_batch = batch;
  return e;
}

namespace
{
struct Bufs
{
  float *a1, *a2, *a3, *out, *g0, *g1;
};

// the layout of the activation buffer for batch size nb (<= act_batch)
Bufs buffers(float *act, int nb, int H, int W)
{
  const size_t cp = (size_t)nb * cnn::kChannels * H * W;
  Bufs b;
  b.a1 = act;
  b.a2 = b.a1 + cp;
  b.a3 = b.a2 + cp;
  b.out = b.a3 + cp;
  b.g0 = b.out + (size_t)nb * H * W;
  b.g1 = b.g0 + cp;
  return b;
}

template <int CIN, int COUT, bool ACT>
void conv(const float *in, const uint32_t *w, const float *bias, float *out,
          int batch, int H, int W, hipStream_t s)
{
  constexpr int OT = COUT < 8 ? COUT : 8;
  hipLaunchKernelGGL((cnn::conv_kernel<CIN, COUT, ACT>),
                     dim3(W / cnn::kTile, H / cnn::kTile,
                          batch * (COUT / OT)),
                     dim3(cnn::kTile * cnn::kTile), 0, s, in, w, bias, out, H,
                     W);
}

template <bool ACT, int OH>
hipError_t conv32_launch(const float *in, const uint32_t *w,
                         const float *bias, float *out, int batch, int H,
                         int W, hipStream_t s)
{
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void *>(&cnn::conv32_mfma_kernel<ACT, OH>),
    hipFuncAttributeMaxDynamicSharedMemorySize, cnn::kConvLds);
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL((cnn::conv32_mfma_kernel<ACT, OH>),
                     dim3(W / cnn::kMTX, H / cnn::kMTY, batch * OH),
                     dim3(256), cnn::kConvLds, s, in, w, bias, out, H, W);
  return hipGetLastError();
}

// small batches split the output channels over two blocks per pixel tile,
// so that a 128 x 128 image still fills the GPU's CUs
template <bool ACT>
hipError_t conv32(const float *in, const uint32_t *w, const float *bias,
                  float *out, int batch, int H, int W, hipStream_t s)
{
  const long long tiles = (long long)batch * (H / cnn::kMTY) * (W / cnn::kMTX);
  return tiles < 512
           ? conv32_launch<ACT, 2>(in, w, bias, out, batch, H, W, s)
           : conv32_launch<ACT, 1>(in, w, bias, out, batch, H, W, s);
}

void conv_to1(const float *in, const uint32_t *w, const float *bias,
              float *out, int batch, int H, int W, hipStream_t s)
{
  hipLaunchKernelGGL(cnn::conv_to1_kernel<cnn::kChannels>,
                     dim3(batch * (W / 16) * (H / 4)), dim3(256), 0, s, in, w,
                     bias, out, H, W);
}

// the weight and bias gradients of a 32 -> 32 layer on the matrix cores
hipError_t wgrad32(const float *in, const float *g, float *part, float *gw,
                   float *gb, int batch, int H, int W, hipStream_t s)
{
  constexpr int C = cnn::kChannels;
  const size_t lds = sizeof(float) * C * (cnn::kGStride + cnn::kInStride);
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void *>(&cnn::wgrad32_mfma_kernel),
    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess)
    return e;
  const int nb = wgrad_blocks(batch, H, W);
  hipLaunchKernelGGL(cnn::wgrad32_mfma_kernel, dim3(nb), dim3(256), lds, s,
                     in, g, part, batch, H, W);
  const int n_out = cnn::kWgradOut;
  // gw and gb are adjacent in the parameter layout (w[Co][Ci][9], b[Co])
  hipLaunchKernelGGL(cnn::part_reduce_kernel, dim3((n_out + 63) / 64),
                     dim3(256), 0, s, part, nb, n_out, C * C * 9, gw, gb);
  return hipGetLastError();
}

// the weight and bias gradients of the 1 -> 32 and 32 -> 1 layers
template <int CIN, int COUT>
hipError_t wgrad_split(const float *in, const float *g, float *part,
                       float *gw, float *gb, int batch, int H, int W,
                       hipStream_t s)
{
  constexpr int pairs = CIN * COUT;
  hipLaunchKernelGGL((cnn::wgrad_split_kernel<CIN, COUT>),
                     dim3(pairs, kSplitChunks), dim3(256), 0, s, in, g, part,
                     batch, H, W);
  hipLaunchKernelGGL(cnn::wgrad_split_reduce_kernel,
                     dim3((pairs * 10 + 255) / 256), dim3(256), 0, s, part,
                     kSplitChunks, CIN, pairs, gw, gb);
  return hipGetLastError();
}

void act_grad(float *g, const float *a, long long n, hipStream_t s)
{
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(cnn::act_grad_kernel, dim3(blocks), dim3(256), 0, s, g, a,
                     n);
}
} // namespace

hipError_t CnnNet::forward(const float *in, float *out, int batch,
                           hipStream_t s)
{
  hipError_t e = reserve(batch);
  if (e != hipSuccess)
    return e;
  const Bufs b = buffers(act, batch, H, W);
  conv<1, 32, true>(in, pk + pack_offset_fwd(0), p + param_offset_b(0), b.a1,
                    batch, H, W, s);
  e = conv32<true>(b.a1, pk + pack_offset_fwd(1), p + param_offset_b(1), b.a2,
                   batch, H, W, s);
  if (e == hipSuccess)
    e = conv32<true>(b.a2, pk + pack_offset_fwd(2), p + param_offset_b(2),
                     b.a3, batch, H, W, s);
  if (e != hipSuccess)
    return e;
  conv_to1(b.a3, pk + pack_offset_fwd(3), p + param_offset_b(3), out, batch,
           H, W, s);
  return hipGetLastError();
}

hipError_t CnnNet::loss_and_grad(const float *in, const float *target,
                                 int batch, hipStream_t s)
{
  hipError_t e = reserve(batch);
  if (e != hipSuccess)
    return e;
  const Bufs b = buffers(act, batch, H, W);
  e = forward(in, b.out, batch, s);
  if (e != hipSuccess)
    return e;
  const long long plane = (long long)H * W;
  const long long nC = (long long)batch * cnn::kChannels * plane;
  const int C = cnn::kChannels;
  // the loss partials sit in the weight-gradient buffer, consumed at once
  double *mse_part = reinterpret_cast<double *>(part);
  hipLaunchKernelGGL(cnn::mse_partial_kernel, dim3(cnn::kMseBlocks),
                     dim3(256), 0, s, b.out, target, b.g0, batch * plane,
                     mse_part);
  hipLaunchKernelGGL(cnn::mse_finish_kernel, dim3(1), dim3(64), 0, s,
                     mse_part, cnn::kMseBlocks, batch * plane, loss_dev);
  // layer 3 (C -> 1, linear)
  e = wgrad_split<32, 1>(b.a3, b.g0, part, grad + param_offset_w(3),
                         grad + param_offset_b(3), batch, H, W, s);
  if (e != hipSuccess)
    return e;
  conv<1, 32, false>(b.g0, pk + pack_offset_bwd(3), nullptr, b.g1, batch, H, W,
                     s);
  act_grad(b.g1, b.a3, nC, s);
  // layer 2
  e = wgrad32(b.a2, b.g1, part, grad + param_offset_w(2),
              grad + param_offset_b(2), batch, H, W, s);
  if (e != hipSuccess)
    return e;
  e = conv32<false>(b.g1, pk + pack_offset_bwd(2), nullptr, b.g0, batch, H, W,
                    s);
  if (e != hipSuccess)
    return e;
  act_grad(b.g0, b.a2, nC, s);
  // layer 1
  e = wgrad32(b.a1, b.g0, part, grad + param_offset_w(1),
              grad + param_offset_b(1), batch, H, W, s);
  if (e != hipSuccess)
    return e;
  e = conv32<false>(b.g0, pk + pack_offset_bwd(1), nullptr, b.g1, batch, H, W,
                    s);
  if (e != hipSuccess)
    return e;
  act_grad(b.g1, b.a1, nC, s);
  // layer 0 (1 -> C)
  return wgrad_split<1, 32>(in, b.g1, part, grad + param_offset_w(0),
                            grad + param_offset_b(0), batch, H, W, s);
}

hipError_t CnnNet::adam(float lr, hipStream_t s)
{
  ++steps;
  const double b1 = 0.9, b2 = 0.999;
  const float c1 = (float)(1.0 - std::pow(b1, steps));
  const float c2 = (float)(1.0 - std::pow(b2, steps));
  const int n = n_params();
  hipLaunchKernelGGL(cnn::adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                     p, m, v, grad, n, lr, (float)b1, (float)b2, c1, c2, 1e-8f);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? pack(s) : e;
}

// ---------------------------------------------------------------------------
// image helpers

hipError_t cnn_dofs_to_image(const double *u, const double *sub, int n0,
                             int n1, int N, float *image, double *scale,
                             hipStream_t s)
{
  const long long n = (long long)n0 * n1 * N * N;
  hipLaunchKernelGGL(cnn::dofs_to_image_kernel, dim3(1), dim3(1024), 0, s, u,
                     sub, n, n0, N, image, scale);
  return hipGetLastError();
}

hipError_t cnn_image_to_dofs(const float *image, const double *scale, int n0,
                             int n1, int N, double *u, hipStream_t s)
{
  const long long n = (long long)n0 * n1 * N * N;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(cnn::image_to_dofs_kernel, dim3(blocks), dim3(256), 0, s,
                     image, scale, n, n0, n1, N, u);
  return hipGetLastError();
}

hipError_t cnn_sample(const float *src, const float *src2, uint64_t seed,
                      uint64_t counter0, int count, int skip, int n,
                      long long plane, float *dst, float *dst2,
                      hipStream_t s)
{
  const long long total = (long long)n * plane;
  const unsigned blocks =
    (unsigned)std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(cnn::gather_kernel, dim3(blocks), dim3(256), 0, s, src,
                     src2, seed, counter0, count, skip, n, plane, dst, dst2);
  return hipGetLastError();
}

uint16_t cnn_to_bf16(float x) { return cnn::to_bf16(x); }
} // namespace hd
