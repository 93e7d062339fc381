// hd_quad_common.h -- device helpers shared by the quadrature-form kernels
// (the element-centric apply in hd_quadrature.hip and the face-centric
// cell / face / inverse-mass kernels in hd_fcl.hip): NQ^D cell arrays in LDS,
// in-place 1D contractions line by line, point geometry and velocity.
#pragma once

#include "hd_device.h"

namespace hd
{
namespace quad
{
constexpr int kBlock = 256;

template <int D, int ND, int NQ>
struct Shape
{
  static constexpr int CELL = ipow(NQ, D);  // quadrature-point array
  static constexpr int CELLD = ipow(ND, D); // nodal values of a cell
  static constexpr int FACE = ipow(NQ, D - 1);
  static constexpr int CPB = CELL >= 1024 ? 1 : 1024 / CELL;
};

// position in the NQ^DD array of the point with base-ND digits i
template <int DD, int ND, int NQ>
__device__ inline int qpos(int i)
{
  if constexpr (ND == NQ)
    return i;
  else
    {
      int r = 0, pw = 1;
#pragma unroll
      for (int j = 0; j < DD; ++j)
        {
          r += (i % ND) * pw;
          i /= ND;
          pw *= NQ;
        }
      return r;
    }
}

// base index of line f (0 <= f < N^{D-1}) along direction d
template <int D, int N>
__device__ inline int line_base(int f, int d)
{
  int base = 0, pw = 1;
#pragma unroll
  for (int j = 0; j < D; ++j)
    {
      if (j != d)
        {
          base += (f % N) * pw;
          f /= N;
        }
      pw *= N;
    }
  return base;
}

// in-place 1D contraction of all lines along d of the CPB cells in buf (NQ^D
// arrays): line[0, NO) <- M line[0, NI), M(r, c) = M[r * sr + c * sc]
template <int D, int NQ, int NI, int NO, class T>
__device__ inline void contract_lines(T *buf, int cellp, int d, const T *M,
                                      int sr, int sc, int ncell)
{
  constexpr int FACE = ipow(NQ, D - 1);
  const int sd = ipow(NQ, d);
  for (int l = threadIdx.x; l < ncell * FACE; l += kBlock)
    {
      const int lc = l / FACE;
      T *p = buf + lc * cellp + line_base<D, NQ>(l - lc * FACE, d);
      T in[NI];
#pragma unroll
      for (int j = 0; j < NI; ++j)
        in[j] = p[j * sd];
#pragma unroll
      for (int i = 0; i < NO; ++i)
        {
          T acc = 0;
#pragma unroll
          for (int j = 0; j < NI; ++j)
            acc += M[i * sr + j * sc] * in[j];
          p[i * sd] = acc;
        }
    }
}

// the 1D tables in the arithmetic type: FP64 straight from the kernel
// arguments, FP32 converted once per block into LDS
template <class T>
struct QTab
{
  const T *S, *Dq, *F, *Sinv;
};

// transport velocity of direction d at the in-cell point pt of `cell`
// (constant, or the Vlasov-Poisson field; a face point's digit d is unused)
template <int D, int N>
__device__ inline double point_velocity(const QuadDev &q, const MeshDev &m,
                                        const double *efield, int dim_x,
                                        int d, long long cell, int pt)
{
  if (!q.vp)
    return q.a[d];
  if (d < dim_x)
    {
      const int e = dim_x + d;
      int pe = 1;
      for (int j = 0; j < e; ++j)
        pe *= N;
      const int qe = (pt / pe) % N;
      const int ce = (int)((cell / m.stride[e]) % m.n[e]);
      return m.lo[e] + (ce + q.qp[qe]) * m.h[e];
    }
  const int e = d - dim_x;
  int nqx = 1;
  for (int j = 0; j < dim_x; ++j)
    nqx *= N;
  const long long xc = cell % q.n_x_cells;
  return efield[(xc * nqx + pt % nqx) * dim_x + e];
}

template <int D, int N>
__device__ inline double jxw_point(const QuadDev &q, const MeshDev &m, int i)
{
  double v = 1.0;
#pragma unroll
  for (int j = 0; j < D; ++j)
    {
      v *= m.h[j] * q.w[i % N];
      i /= N;
    }
  return v;
}

} // namespace quad
} // namespace hd
