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
  // cells per block: about 2048 points, at most one cell per thread (the
  // per-cell set-up uses thread c for cell c)
  static constexpr int CPB = CELL >= 2048        ? 1
                             : 2048 / CELL > kBlock ? kBlock
                                                    : 2048 / CELL;
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

// a cell's coordinates and the indices of its x- and v-space cells (cell =
// xc + n_x vc), computed once per cell so that the per-point helpers below
// do no 64-bit division
struct CellAt
{
  int c[kMaxDim];
  long long xc, vc;
};

__device__ inline CellAt cell_at(const MeshDev &m, long long cell,
                                 long long n_x_cells)
{
  CellAt a;
  long long r = cell;
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j)
    if (j < m.dim)
      {
        a.c[j] = (int)(r % m.n[j]);
        r /= m.n[j];
      }
    else
      a.c[j] = 0;
  a.xc = n_x_cells > 0 ? cell % n_x_cells : 0;
  a.vc = n_x_cells > 0 ? cell / n_x_cells : 0;
  return a;
}

// transport velocity of direction d at the in-cell point pt of a cell
// (constant, or the Vlasov-Poisson field; a face point's digit d is unused)
template <int D, int N>
__device__ inline double point_velocity_at(const QuadDev &q, const MeshDev &m,
                                           const double *efield, int dim_x,
                                           int d, const CellAt &ca, int pt)
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
      return m.lo[e] + (ca.c[e] + q.qp[qe]) * m.h[e];
    }
  const int e = d - dim_x;
  int nqx = 1;
  for (int j = 0; j < dim_x; ++j)
    nqx *= N;
  return efield[(ca.xc * nqx + pt % nqx) * dim_x + e];
}

template <int D, int N>
__device__ inline double point_velocity(const QuadDev &q, const MeshDev &m,
                                        const double *efield, int dim_x,
                                        int d, long long cell, int pt)
{
  return point_velocity_at<D, N>(q, m, efield, dim_x, d,
                                 cell_at(m, cell, q.vp ? q.n_x_cells : 0), pt);
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

// ---------------------------------------------------------------------------
// deformed meshes: the per-point factors from the per-space tables of
// DeformDev (the Jacobian is block diagonal over x and v, so every factor is
// an x-space value times a v-space value)

// x and v parts of a point index of the NQ^D cell array
template <int NQ>
__device__ inline void split_point(const DeformDev &g, int pt, int &px,
                                   int &pv, int &nqx, int &nqv)
{
  nqx = 1;
  nqv = 1;
  for (int j = 0; j < g.dim_x; ++j)
    nqx *= NQ;
  for (int j = 0; j < g.dim_v; ++j)
    nqv *= NQ;
  px = pt % nqx;
  pv = pt / nqx;
}

// |det J| w at cell point pt: the JxW of the deformed cell (x / v cells cx,
// cv)
template <int NQ>
__device__ inline double deformed_jxw_at(const DeformDev &g, long long cx,
                                         long long cv, int pt)
{
  int px, pv, nqx, nqv;
  split_point<NQ>(g, pt, px, pv, nqx, nqv);
  return g.xca[(cx * nqx + px) * (g.dim_x + 1) + g.dim_x] *
         g.vca[(cv * nqv + pv) * (g.dim_v + 1) + g.dim_v];
}

template <int NQ>
__device__ inline double deformed_jxw(const DeformDev &g, long long cell,
                                      int pt)
{
  return deformed_jxw_at<NQ>(g, cell % g.n_x_cells, cell / g.n_x_cells, pt);
}

// (J^{-1} a)_d JxW at cell point pt (direction d of phase space)
template <int NQ>
__device__ inline double deformed_velocity_jxw_at(const DeformDev &g,
                                                  long long cx, long long cv,
                                                  int pt, int d)
{
  int px, pv, nqx, nqv;
  split_point<NQ>(g, pt, px, pv, nqx, nqv);
  const double *xc = g.xca + (cx * nqx + px) * (g.dim_x + 1);
  const double *vc = g.vca + (cv * nqv + pv) * (g.dim_v + 1);
  const double jw = xc[g.dim_x] * vc[g.dim_v];
  return (d < g.dim_x ? xc[d] : vc[d - g.dim_x]) * jw;
}

template <int NQ>
__device__ inline double deformed_velocity_jxw(const DeformDev &g,
                                               long long cell, int pt, int d)
{
  return deformed_velocity_jxw_at<NQ>(g, cell % g.n_x_cells,
                                      cell / g.n_x_cells, pt, d);
}

// (a . n) JxW_face at the face point of the cell array position `base`
// (digit d unused) of the face of direction d, side `side`
template <int NQ>
__device__ inline double deformed_face_at(const DeformDev &g, long long cx,
                                          long long cv, int base, int d,
                                          int side)
{
  int px, pv, nqx, nqv;
  split_point<NQ>(g, base, px, pv, nqx, nqv);
  const bool in_x = d < g.dim_x;
  const int dl = in_x ? d : d - g.dim_x;      // direction within its space
  const int dim = in_x ? g.dim_x : g.dim_v;
  int p = in_x ? px : pv;
  // the face point: the space's point digits without digit dl
  int fp = 0, pw = 1;
  for (int j = 0; j < dim; ++j)
    {
      const int dg = p % NQ;
      p /= NQ;
      if (j != dl)
        {
          fp += dg * pw;
          pw *= NQ;
        }
    }
  const int nfp = (in_x ? nqx : nqv) / NQ;
  const long long c = in_x ? cx : cv;
  const double gf =
    (in_x ? g.xfg : g.vfg)[(c * 2 * dim + 2 * dl + side) * nfp + fp];
  // the other space's JxW at its point
  const double jw = in_x ? g.vca[(cv * nqv + pv) * (g.dim_v + 1) + g.dim_v]
                         : g.xca[(cx * nqx + px) * (g.dim_x + 1) + g.dim_x];
  return gf * jw;
}

template <int NQ>
__device__ inline double deformed_face(const DeformDev &g, long long cell,
                                       int base, int d, int side)
{
  return deformed_face_at<NQ>(g, cell % g.n_x_cells, cell / g.n_x_cells, base,
                              d, side);
}
} // namespace quad
} // namespace hd
