// hd_util.hip -- vector kernels around the operator: interpolation of the
// initial data (numerics/vector_tools.h:81-135), the seeded benchmark noise,
// L2 norm and error at the quadrature points (vector_tools.h:144-223) and
// precision conversion.
#include "hd_device.h"
#include "hd_functions.h"

#include <cmath>

namespace hd
{
namespace util
{

__global__ void interpolate_kernel(double *dst, MeshDev mesh, FuncDev f,
                                   double t, long long n_dofs, int dofs_cell,
                                   DeformDev geo)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_dofs)
    return;
  long long cell = i / dofs_cell;
  int node = (int)(i - cell * dofs_cell);
  double x[kMaxDim];
  for (int d = 0; d < mesh.dim; ++d)
    {
      const int c = (int)(cell % mesh.n[d]);
      cell /= mesh.n[d];
      const int j = node % f.nd;
      node /= f.nd;
      x[d] = mesh.lo[d] + (c + f.nodes[j]) * mesh.h[d];
    }
  if (geo.on)
    deform_point(geo, mesh, x);
  dst[i] = eval_function(f.id, f.prm, f.dim_x, mesh.dim, x, t);
}

__global__ void noise_kernel(double *dst, uint64_t seed, double amplitude,
                             long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    dst[i] += amplitude * splitmix_uniform(seed, (uint64_t)i);
}

template <typename T>
__global__ void convert_kernel(const double *src, T *dst, long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    dst[i] = (T)src[i];
}

// Blocks stride over the cells (one cell per block when the grid has a block
// per cell): interpolate to the quadrature points with S^D through two LDS
// buffers (ping-pong per direction), accumulate f^2 JxW and (f - g)^2 JxW
// per thread, then block-reduce them into the block's two partial sums.
__global__ __launch_bounds__(256) void norm_kernel(const double *src,
                                                   MeshDev mesh, NormDev nt,
                                                   FuncDev f, double t,
                                                   int vol, double *partial,
                                                   DeformDev geo)
{
  extern __shared__ double sm[];
  const int D = mesh.dim, nd = nt.nd, nq = nt.nq;
  double *b0 = sm, *b1 = sm + vol, *red = sm + 2 * vol;
  int dofs = 1, qvol = 1;
  for (int d = 0; d < D; ++d)
    {
      dofs *= nd;
      qvol *= nq;
    }
  double s0 = 0.0, s1 = 0.0;
  for (long long cell = blockIdx.x; cell < mesh.n_cells; cell += gridDim.x)
    {
      __syncthreads(); // the previous cell's LDS reads are done
      for (int i = threadIdx.x; i < dofs; i += blockDim.x)
        b0[i] = src[cell * dofs + i];
      __syncthreads();
      double *in = b0, *out = b1;
      for (int d = 0; d < D; ++d)
        {
          // in: extents nq (j < d), nd (j >= d); out: nq along d as well
          int pre = 1, post = 1;
          for (int j = 0; j < d; ++j)
            pre *= nq;
          for (int j = d + 1; j < D; ++j)
            post *= nd;
          const int n_out = pre * nq * post;
          for (int o = threadIdx.x; o < n_out; o += blockDim.x)
            {
              const int pi = o % pre, rest = o / pre;
              const int q = rest % nq, po = rest / nq;
              double acc = 0.0;
              for (int j = 0; j < nd; ++j)
                acc += nt.S[q * nd + j] * in[pi + pre * (j + nd * po)];
              out[o] = acc;
            }
          __syncthreads();
          double *tmp = in;
          in = out;
          out = tmp;
        }
      for (int q = threadIdx.x; q < qvol; q += blockDim.x)
        {
          double x[kMaxDim], jxw = 1.0;
          long long c = cell;
          int r = q;
          for (int d = 0; d < D; ++d)
            {
              const int cd = (int)(c % mesh.n[d]);
              c /= mesh.n[d];
              const int qd = r % nq;
              r /= nq;
              x[d] = mesh.lo[d] + (cd + nt.qp[qd]) * mesh.h[d];
              jxw *= mesh.h[d] * nt.w[qd];
            }
          if (geo.on)
            {
              // the mapped point; JxW = (|det J_x| w_x)(|det J_v| w_v)
              deform_point(geo, mesh, x);
              int nqx = 1;
              for (int d = 0; d < geo.dim_x; ++d)
                nqx *= nq;
              const int nqv = qvol / nqx;
              const long long cx = cell % geo.n_x_cells;
              const long long cv = cell / geo.n_x_cells;
              jxw = geo.xjw[cx * nqx + q % nqx] * geo.vjw[cv * nqv + q / nqx];
            }
          const double v = in[q];
          const double e = v - eval_function(f.id, f.prm, f.dim_x, D, x, t);
          s0 += v * v * jxw;
          s1 += e * e * jxw;
        }
    }
  red[threadIdx.x] = s0;
  red[blockDim.x + threadIdx.x] = s1;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s /= 2)
    {
      if ((int)threadIdx.x < s)
        {
          red[threadIdx.x] += red[threadIdx.x + s];
          red[blockDim.x + threadIdx.x] += red[blockDim.x + threadIdx.x + s];
        }
      __syncthreads();
    }
  if (threadIdx.x == 0)
    {
      partial[2 * blockIdx.x] = red[0];
      partial[2 * blockIdx.x + 1] = red[blockDim.x];
    }
}

// out = (sqrt(sum of the n block partials of f^2 JxW), sqrt(... of the
// error)); one block, a fixed summation order
__global__ __launch_bounds__(256) void norm_finish_kernel(
  const double *__restrict__ partial, int n, double *__restrict__ out)
{
  __shared__ double red[512];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    {
      s0 += partial[2 * i];
      s1 += partial[2 * i + 1];
    }
  red[threadIdx.x] = s0;
  red[256 + threadIdx.x] = s1;
  __syncthreads();
  for (int s = 128; s > 0; s /= 2)
    {
      if ((int)threadIdx.x < s)
        {
          red[threadIdx.x] += red[threadIdx.x + s];
          red[256 + threadIdx.x] += red[256 + threadIdx.x + s];
        }
      __syncthreads();
    }
  if (threadIdx.x == 0)
    {
      out[0] = sqrt(red[0]);
      out[1] = sqrt(red[256]);
    }
}
} // namespace util

hipError_t launch_interpolate(double *dst, const MeshDev &mesh,
                              const FuncDev &f, double t, long long n_dofs,
                              int dofs_cell, const DeformDev &geo,
                              hipStream_t stream)
{
  hipLaunchKernelGGL(util::interpolate_kernel,
                     dim3((unsigned)((n_dofs + 255) / 256)), dim3(256), 0,
                     stream, dst, mesh, f, t, n_dofs, dofs_cell, geo);
  return hipGetLastError();
}

hipError_t launch_noise(double *dst, uint64_t seed, double amplitude,
                        long long n, hipStream_t stream)
{
  hipLaunchKernelGGL(util::noise_kernel, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, stream, dst, seed, amplitude, n);
  return hipGetLastError();
}

hipError_t launch_convert_f32(const double *src, float *dst, long long n,
                              hipStream_t stream)
{
  hipLaunchKernelGGL(util::convert_kernel<float>,
                     dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     src, dst, n);
  return hipGetLastError();
}

hipError_t launch_norm(const double *src, const MeshDev &mesh,
                       const NormDev &nt, const FuncDev &f, double t, int vol,
                       double *partial, const DeformDev &geo,
                       hipStream_t stream, int blocks)
{
  const size_t lds = (2 * (size_t)vol + 512) * sizeof(double);
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void *>(&util::norm_kernel),
    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess)
    return e;
  const unsigned grid = blocks > 0 ? (unsigned)blocks
                                   : (unsigned)mesh.n_cells;
  hipLaunchKernelGGL(util::norm_kernel, dim3(grid), dim3(256), lds, stream,
                     src, mesh, nt, f, t, vol, partial, geo);
  return hipGetLastError();
}

hipError_t launch_norm_finish(const double *partial, int n, double *out,
                              hipStream_t stream)
{
  hipLaunchKernelGGL(util::norm_finish_kernel, dim3(1), dim3(256), 0, stream,
                     partial, n, out);
  return hipGetLastError();
}
} // namespace hd
