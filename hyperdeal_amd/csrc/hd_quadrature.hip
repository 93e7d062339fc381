// hd_quadrature.hip -- the general ECL apply of the phase-space DG advection
// operator on gfx950: the reference's quadrature algorithm step by step
// (operators/advection/advection_operation.h:221-566), for Cartesian meshes
// with periodic or Dirichlet boundaries, any skew factor, Gauss quadrature
// with k+1 or k+2 points, or Gauss-Lobatto collocation.
//
// ND = k+1 nodes and NQ >= ND quadrature points per direction.  One block
// works on CPB cells held in LDS as NQ^D arrays (the nodal values sit at the
// positions whose digits are all < ND):
//   U  <- src cell (nodal), then U <- S^D U             (quadrature values)
//   R  <- cell integral: -skew (a.grad u) JxW + sum_d D_d^T[(1-skew) u a_d/h_d JxW]
//   per direction d, its two faces s together: u- = F_s . U along d;
//       u+ = S^{D-1} (neighbour's face nodes) or the Dirichlet ghost
//       -u- + 2g; R += F_s^T (-flux JxW_f)
//   R <- Sinv^D (R / JxW)                                 (inverse mass)
// S is NQ x ND, Sinv = (S^T W S)^-1 S^T W is ND x NQ.  Every 1D contraction
// is done line by line: a thread owns whole lines along the current
// direction (reading ND or NQ values, writing NQ or ND), so updates are in
// place and only a barrier separates consecutive directions.  Lines through
// positions that hold no value yet (digits >= ND of a direction not yet
// expanded, or already reduced) compute values that are never read.
#include "hd_device.h"
#include "hd_functions.h"
#include "hd_quad_common.h"

#include <cmath>

namespace hd
{
namespace quad
{
// MODE 0: dst = apply(src); MODE 1: LSRK stage epilogue
// T: vector, LDS and contraction type (double, or float for the FP32-mixed
// form: the LSRK update of the FP64 solution stays in FP64)
// Three waves per SIMD (at most 168 registers): the FP32 2+2 k=3 kernel
// needs 170 without the bound and runs at two waves 25% slower; the
// spills the bound causes elsewhere (12-80 B per lane) cost less.
template <int D, int ND, int NQ, int MODE, class T>
__global__ __launch_bounds__(kBlock, 3) void apply_kernel(
  const T *__restrict__ src, T *__restrict__ dst, QuadDev q, MeshDev mesh,
  StageDev stage, double time, int dim_x, int collocation,
  const double *__restrict__ efield, DeformDev geo)
{
  using Sh = Shape<D, ND, NQ>;
  constexpr int CELL = Sh::CELL, CELLD = Sh::CELLD, FACE = Sh::FACE;
  constexpr int CPB = Sh::CPB;
  constexpr int CELLP = CELL + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *U = reinterpret_cast<T *>(smem_raw); // CPB x CELLP
  T *R = U + CPB * CELLP;                 // CPB x CELLP
  T *Fp = R + CPB * CELLP;                // 2 sides x CPB x FACE (u+)
  QTab<T> tb;
  if constexpr (sizeof(T) == sizeof(double))
    tb = {q.S, q.Dq, q.F, q.Sinv};
  else
    {
      __shared__ T tabs[64 + 64 + 16 + 64];
      for (int i = threadIdx.x; i < 64; i += kBlock)
        {
          tabs[i] = (T)q.S[i];
          tabs[64 + i] = (T)q.Dq[i];
          tabs[144 + i] = (T)q.Sinv[i];
          if (i < 16)
            tabs[128 + i] = (T)q.F[i];
        }
      tb = {tabs, tabs + 64, tabs + 128, tabs + 144};
    }

  const long long cell0 = (long long)blockIdx.x * CPB;
  const int ncell = (int)min((long long)CPB, mesh.n_cells - cell0);
  const int tid = threadIdx.x;
  // the block's cells: coordinates and x / v cells, once (read after the
  // barrier that follows the load of U)
  __shared__ CellAt at[CPB];
  if (tid < ncell)
    at[tid] = cell_at(mesh, cell0 + tid,
                      q.vp ? q.n_x_cells : (geo.on ? geo.n_x_cells : 0));

  // 1) nodal -> quadrature values (S: NQ x ND).  The load does direction 0
  // itself: a thread owns a line along direction 0 of U, reads the line's
  // ND nodal values (contiguous in the cell) and writes its NQ points;
  // lines through positions that hold no nodal value (a digit >= ND in
  // another direction) are zero.
  if (!collocation)
    {
      for (int l = tid; l < ncell * FACE; l += kBlock)
        {
          const int lc = l / FACE, fi = l - lc * FACE;
          int rem = fi, nb = 0, pwn = ND;
          bool node = true;
#pragma unroll
          for (int j = 1; j < D; ++j)
            {
              const int dg = rem % NQ;
              rem /= NQ;
              node = node && dg < ND;
              nb += dg * pwn;
              pwn *= ND;
            }
          T in[ND];
#pragma unroll
          for (int t = 0; t < ND; ++t)
            in[t] = node ? src[(cell0 + lc) * CELLD + nb + t] : T(0);
          T *pl = U + lc * CELLP + fi * NQ;
#pragma unroll
          for (int i = 0; i < NQ; ++i)
            {
              T acc = 0;
#pragma unroll
              for (int t = 0; t < ND; ++t)
                acc += tb.S[i * ND + t] * in[t];
              pl[i] = acc;
            }
        }
      __syncthreads();
      for (int d = 1; d < D; ++d)
        {
          contract_lines<D, NQ, ND, NQ, T>(U, CELLP, d, tb.S, ND, 1, ncell);
          __syncthreads();
        }
    }
  else
    {
      // collocation: the nodes are the points (ND = NQ)
      for (int i = tid; i < ncell * CELLD; i += kBlock)
        {
          const int lc = i / CELLD;
          U[lc * CELLP + (i - lc * CELLD)] = src[cell0 * CELLD + i];
        }
      __syncthreads();
    }

  // 2) cell integral into R (direction 0 writes R, the others add)
  for (int d = 0; d < D; ++d)
    {
      const int sd = ipow(NQ, d);
      for (int l = tid; l < ncell * FACE; l += kBlock)
        {
          const int lc = l / FACE;
          const int base = line_base<D, NQ>(l - lc * FACE, d);
          const T *u = U + lc * CELLP + base;
          T *r = R + lc * CELLP + base;
          T uq[NQ], jw[NQ], acc[NQ];
          const T skew = (T)q.skew;
#pragma unroll
          for (int j = 0; j < NQ; ++j)
            {
              uq[j] = u[j * sd];
              // velocity / h folded into the quadrature weight (deformed:
              // the contravariant velocity (J^{-1} a)_d times JxW)
              jw[j] = (T)(geo.on ? deformed_velocity_jxw_at<NQ>(
                                     geo, at[lc].xc, at[lc].vc, base + j * sd,
                                     d)
                                 : jxw_point<D, NQ>(q, mesh, base + j * sd) *
                                     point_velocity_at<D, NQ>(
                                       q, mesh, efield, dim_x, d, at[lc],
                                       base + j * sd) /
                                     mesh.h[d]);
              acc[j] = 0;
            }
          if (q.skew != 0.0)
#pragma unroll
            for (int i = 0; i < NQ; ++i)
              {
                T g = 0;
#pragma unroll
                for (int j = 0; j < NQ; ++j)
                  g += tb.Dq[i * NQ + j] * uq[j];
                acc[i] += -skew * g * jw[i];
              }
          if (q.skew != 1.0)
#pragma unroll
            for (int i = 0; i < NQ; ++i)
              {
                T s = 0;
#pragma unroll
                for (int j = 0; j < NQ; ++j)
                  s += tb.Dq[j * NQ + i] * ((1 - skew) * uq[j] * jw[j]);
                acc[i] += s;
              }
#pragma unroll
          for (int i = 0; i < NQ; ++i)
            r[i * sd] = d == 0 ? acc[i] : r[i * sd] + acc[i];
        }
      __syncthreads();
    }

  // 3) faces, both sides of a direction at once (side 0: the face towards
  // the left neighbour, normal -e_d; side 1: towards the right one).  The
  // lifts of the two sides go into R in the order side 0, side 1, as if the
  // faces were taken one by one.
  for (int d = 0; d < D; ++d)
    {
      const int sd = ipow(NQ, d);
      const int sdn = ipow(ND, d); // stride of direction d in the src cell
      const long long s = mesh.stride[d];
      const long long wrap = (long long)(mesh.n[d] - 1) * s;
      // the neighbours' face nodes (they sit at the face positions whose
      // digits are all < ND; the others are zero) at the face points of
      // both sides: u+ = S^{D-1} (face nodes), S: NQ x ND.  The gather does
      // the first face direction j0 itself (a thread owns a line along j0
      // and reads the line's ND nodes from the neighbour), the passes after
      // it the others.
      const int j0 = d == 0 ? 1 : 0;
      if (!collocation)
        {
          const int sn0 = ipow(ND, j0); // node stride of j0 in the src cell
          const int nl = FACE / NQ;
          for (int l = tid; l < 2 * ncell * nl; l += kBlock)
            {
              const int side = l >= ncell * nl;
              const int ls = l - side * ncell * nl;
              const int lc = ls / nl;
              // the line's digits in the face directions after j0: its
              // base in the face array and its node offset in the
              // neighbour
              int rem = ls - lc * nl, fb = 0, nb = 0, pwq = 1, pwn = 1;
              bool node = true;
#pragma unroll
              for (int j = 0; j < D; ++j)
                {
                  if (j != d)
                    {
                      if (j != j0)
                        {
                          const int dg = rem % NQ;
                          rem /= NQ;
                          node = node && dg < ND;
                          fb += dg * pwq;
                          nb += dg * pwn;
                        }
                      pwq *= NQ;
                    }
                  pwn *= ND;
                }
              const long long cell = cell0 + lc;
              const int c = at[lc].c[d];
              long long nc = -1;
              int off;
              if (side == 0)
                {
                  // the left neighbour's last layer
                  off = nb + (ND - 1) * sdn;
                  if (c - 1 >= 0)
                    nc = cell - s;
                  else if (q.periodic[d])
                    nc = cell + wrap;
                }
              else
                {
                  // the right neighbour's first layer
                  off = nb;
                  if (c + 1 < mesh.n[d])
                    nc = cell + s;
                  else if (q.periodic[d])
                    nc = cell - wrap;
                }
              T in[ND];
#pragma unroll
              for (int t = 0; t < ND; ++t)
                in[t] = node && nc >= 0 ? src[nc * CELLD + off + t * sn0]
                                        : T(0);
              T *pl = Fp + side * CPB * FACE + lc * FACE + fb;
#pragma unroll
              for (int i = 0; i < NQ; ++i)
                {
                  T acc = 0;
#pragma unroll
                  for (int t = 0; t < ND; ++t)
                    acc += tb.S[i * ND + t] * in[t];
                  pl[i] = acc;
                }
            }
          __syncthreads();
          for (int j = 0; j < D; ++j)
            {
              if (j == d || j == j0)
                continue;
              // lines of the face arrays along j (a face array is the cell
              // array with direction d of extent 1)
              const int jj = j < d ? j : j - 1;
              const int sj = ipow(NQ, jj);
              for (int l = tid; l < 2 * ncell * nl; l += kBlock)
                {
                  const int side = l >= ncell * nl;
                  const int ls = l - side * ncell * nl;
                  const int lc = ls / nl;
                  int li = ls - lc * nl;
                  // base of line li along face-direction jj in NQ^{D-1}
                  int base = 0, pw = 1;
                  for (int k = 0; k < D - 1; ++k)
                    {
                      if (k != jj)
                        {
                          base += (li % NQ) * pw;
                          li /= NQ;
                        }
                      pw *= NQ;
                    }
                  T *p = Fp + side * CPB * FACE + lc * FACE + base;
                  T in[ND];
#pragma unroll
                  for (int t = 0; t < ND; ++t)
                    in[t] = p[t * sj];
#pragma unroll
                  for (int i = 0; i < NQ; ++i)
                    {
                      T acc = 0;
#pragma unroll
                      for (int t = 0; t < ND; ++t)
                        acc += tb.S[i * ND + t] * in[t];
                      p[i * sj] = acc;
                    }
                }
              __syncthreads();
            }
        }
      else
        {
          // collocation: the face nodes are the face points
          for (int l = tid; l < ncell * FACE; l += kBlock)
            {
              const int lc = l / FACE, fi = l - lc * FACE;
              int nbase = 0;
              {
                int rem = fi, pw = 1;
#pragma unroll
                for (int j = 0; j < D; ++j)
                  {
                    if (j != d)
                      {
                        nbase += (rem % NQ) * pw;
                        rem /= NQ;
                      }
                    pw *= ND;
                  }
              }
              const long long cell = cell0 + lc;
              const int c = at[lc].c[d];
              T up0 = 0, up1 = 0;
              const int off0 = nbase + (ND - 1) * sdn;
              if (c - 1 >= 0)
                up0 = src[(cell - s) * CELLD + off0];
              else if (q.periodic[d])
                up0 = src[(cell + wrap) * CELLD + off0];
              if (c + 1 < mesh.n[d])
                up1 = src[(cell + s) * CELLD + nbase];
              else if (q.periodic[d])
                up1 = src[(cell - wrap) * CELLD + nbase];
              Fp[l] = up0;
              Fp[CPB * FACE + l] = up1;
            }
          __syncthreads();
        }
      // fluxes at the face points of both sides, then their lifts into R
      for (int l = tid; l < ncell * FACE; l += kBlock)
        {
          const int lc = l / FACE, fi = l - lc * FACE;
          const int base = line_base<D, NQ>(fi, d);
          const int c = at[lc].c[d];
          // u- of both sides: the traces of the line along d
          T umt[2] = {0, 0};
          {
            const T *u = U + lc * CELLP + base;
#pragma unroll
            for (int p = 0; p < NQ; ++p)
              {
                const T v = u[p * sd];
                umt[0] += tb.F[p] * v;
                umt[1] += tb.F[NQ + p] * v;
              }
          }
          // face JxW: product over j != d
          double jf = 1.0;
          {
            int rem = base;
#pragma unroll
            for (int j = 0; j < D; ++j)
              {
                if (j != d)
                  jf *= mesh.h[j] * q.w[rem % NQ];
                rem /= NQ;
              }
          }
          // Cartesian: the velocity component at the face point (the same
          // for both sides; the normal flips its sign)
          const double a_pt =
            geo.on ? 0.0
                   : point_velocity_at<D, NQ>(q, mesh, efield, dim_x, d,
                                              at[lc], base);
          T fd[2];
#pragma unroll
          for (int side = 0; side < 2; ++side)
            {
              const bool boundary =
                !q.periodic[d] && (side ? c == mesh.n[d] - 1 : c == 0);
              const double um = umt[side];
              double up = Fp[side * CPB * FACE + l];
              double jfs = jf;
              if (boundary)
                {
                  double g = 0.0;
                  if (q.bc_inhomogeneous)
                    {
                      double x[kMaxDim];
                      int rem = base;
                      for (int j = 0; j < D; ++j)
                        {
                          const int cj = at[lc].c[j];
                          const int ij = rem % NQ;
                          rem /= NQ;
                          x[j] = mesh.lo[j] +
                                 (cj + (j == d ? (double)side : q.qp[ij])) *
                                   mesh.h[j];
                        }
                      if (geo.on)
                        deform_point(geo, mesh, x);
                      g = eval_function(q.bc_function, q.bc_params, dim_x, D,
                                        x, time);
                    }
                  up = -um + 2.0 * g;
                }
              // deformed: (a . n) JxW_face from the tables (the face JxW is
              // in it)
              double an;
              if (geo.on)
                {
                  an = deformed_face_at<NQ>(geo, at[lc].xc, at[lc].vc, base,
                                            d, side);
                  jfs = 1.0;
                }
              else
                an = side ? a_pt : -a_pt;
              const double flux =
                0.5 * ((um + up) * an + fabs(an) * (um - up));
              fd[side] = (T)(-(flux - q.skew * um * an) * jfs);
            }
          T *r = R + lc * CELLP + base;
#pragma unroll
          for (int p = 0; p < NQ; ++p)
            {
              T v = r[p * sd];
              v += tb.F[p] * fd[0];
              v += tb.F[NQ + p] * fd[1];
              r[p * sd] = v;
            }
        }
      __syncthreads();
    }

  // 4) inverse mass matrix (Sinv: ND x NQ)
  for (int i = tid; i < ncell * CELL; i += kBlock)
    {
      const int lc = i / CELL, li = i - lc * CELL;
      R[lc * CELLP + li] /=
        (T)(geo.on ? deformed_jxw_at<NQ>(geo, at[lc].xc, at[lc].vc, li)
                   : jxw_point<D, NQ>(q, mesh, li));
    }
  __syncthreads();
  // the result of a point: the apply's value or the LSRK stage update
  auto store = [&](long long g, T k) {
    if (MODE == 0)
      dst[g] = k;
    else
      {
        const double s = stage.sol[g];
        stage.sol[g] = s + stage.b_dt * (double)k;
        if (stage.ti_out)
          static_cast<T *>(stage.ti_out)[g] = (T)(s + stage.a_dt * (double)k);
      }
  };
  if (!collocation)
    {
      for (int d = D - 1; d >= 1; --d)
        {
          contract_lines<D, NQ, NQ, ND, T>(R, CELLP, d, tb.Sinv, NQ, 1,
                                           ncell);
          __syncthreads();
        }
      // 5) direction 0 and the write: a thread owns a line along direction
      // 0 whose other digits are all < ND and stores its ND nodal values
      // (contiguous in the cell)
      for (int l = tid; l < ncell * FACE; l += kBlock)
        {
          const int lc = l / FACE, fi = l - lc * FACE;
          int rem = fi, nb = 0, pwn = ND;
          bool node = true;
#pragma unroll
          for (int j = 1; j < D; ++j)
            {
              const int dg = rem % NQ;
              rem /= NQ;
              node = node && dg < ND;
              nb += dg * pwn;
              pwn *= ND;
            }
          if (!node)
            continue;
          const T *pl = R + lc * CELLP + fi * NQ;
          T v[NQ];
#pragma unroll
          for (int j = 0; j < NQ; ++j)
            v[j] = pl[j];
          const long long g0 = (cell0 + lc) * CELLD + nb;
#pragma unroll
          for (int i = 0; i < ND; ++i)
            {
              T acc = 0;
#pragma unroll
              for (int j = 0; j < NQ; ++j)
                acc += tb.Sinv[i * NQ + j] * v[j];
              store(g0 + i, acc);
            }
        }
    }
  else
    // 5) write (collocation: the points are the nodes)
    for (int i = tid; i < ncell * CELLD; i += kBlock)
      {
        const int lc = i / CELLD;
        store(cell0 * CELLD + i, R[lc * CELLP + (i - lc * CELLD)]);
      }
}

template <int D, int ND, int NQ, class T>
hipError_t launch(int mode, const T *src, T *dst, const QuadDev &q,
                  const MeshDev &mesh, const StageDev &stage, double time,
                  int dim_x, int collocation, const double *efield,
                  const DeformDev &geo, hipStream_t stream)
{
  using Sh = Shape<D, ND, NQ>;
  const size_t lds = sizeof(T) *
                     (2 * Sh::CPB * (Sh::CELL + 1) + 2 * Sh::CPB * Sh::FACE);
  const long long grid = (mesh.n_cells + Sh::CPB - 1) / Sh::CPB;
  // gfx950 has 160 KB of LDS per CU; opt in above the 64 KB default
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void *>(mode == 0
                                     ? &apply_kernel<D, ND, NQ, 0, T>
                                     : &apply_kernel<D, ND, NQ, 1, T>),
    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess)
    return e;
  if (mode == 0)
    hipLaunchKernelGGL((apply_kernel<D, ND, NQ, 0, T>), dim3((unsigned)grid),
                       dim3(kBlock), lds, stream, src, dst, q, mesh, stage,
                       time, dim_x, collocation, efield, geo);
  else
    hipLaunchKernelGGL((apply_kernel<D, ND, NQ, 1, T>), dim3((unsigned)grid),
                       dim3(kBlock), lds, stream, src, dst, q, mesh, stage,
                       time, dim_x, collocation, efield, geo);
  return hipGetLastError();
}
} // namespace quad

namespace
{
// the instantiated (D, ND, NQ): NQ = ND (k+1 points) or ND + 1 (k+2)
template <class T>
hipError_t dispatch(int D, int ND, int NQ, int mode, const T *src, T *dst,
                    const QuadDev &q, const MeshDev &mesh,
                    const StageDev &stage, double time, int dim_x,
                    int collocation, const double *efield,
                    const DeformDev &geo, hipStream_t stream)
{
#define HD_CASE(DD, ND_, NQ_)                                                \
  if (D == DD && ND == ND_ && NQ == NQ_)                                     \
    return quad::launch<DD, ND_, NQ_, T>(mode, src, dst, q, mesh, stage,     \
                                         time, dim_x, collocation, efield,   \
                                         geo, stream);
#define HD_CASE2(DD, ND_) HD_CASE(DD, ND_, ND_) HD_CASE(DD, ND_, ND_ + 1)
  HD_CASE2(2, 2) HD_CASE2(2, 3) HD_CASE2(2, 4) HD_CASE2(2, 5) HD_CASE(2, 6, 6)
  HD_CASE2(3, 2) HD_CASE2(3, 3) HD_CASE2(3, 4) HD_CASE2(3, 5) HD_CASE(3, 6, 6)
  HD_CASE2(4, 2) HD_CASE2(4, 3) HD_CASE2(4, 4) HD_CASE2(4, 5) HD_CASE(4, 6, 6)
  HD_CASE2(5, 2) HD_CASE2(5, 3) HD_CASE2(5, 4)
  HD_CASE2(6, 2) HD_CASE2(6, 3) HD_CASE(6, 4, 4)
#undef HD_CASE2
#undef HD_CASE
  return hipErrorInvalidValue;
}
} // namespace

bool quadrature_supported(int D, int ND, int NQ)
{
  if (D < 2 || D > 6 || ND < 2 || (NQ != ND && NQ != ND + 1))
    return false;
  // the instantiated node counts: k+1 points up to ND = 6 (D <= 4) or 4;
  // k+2 points up to ND = 5 (D <= 4), 4 (D = 5) or 3 (D = 6)
  const int max_nd =
    NQ == ND ? (D <= 4 ? 6 : 4) : (D <= 4 ? 5 : D == 5 ? 4 : 3);
  if (ND > max_nd)
    return false;
  // both LDS cell buffers and the u+ face buffers of two sides must fit in
  // 160 KB
  const long long cell = (long long)std::pow(NQ, D);
  return 2 * (cell + 1) * 8 + 2 * (cell / NQ) * 8 <= 160 * 1024;
}

hipError_t quadrature_apply(int D, int ND, int NQ, int mode,
                            const double *src, double *dst, const QuadDev &q,
                            const MeshDev &mesh, const StageDev &stage,
                            double time, int dim_x, int collocation,
                            const double *efield, const DeformDev &geo,
                            hipStream_t stream)
{
  return dispatch<double>(D, ND, NQ, mode, src, dst, q, mesh, stage, time,
                          dim_x, collocation, efield, geo, stream);
}

hipError_t quadrature_apply_f32(int D, int ND, int NQ, int mode,
                                const float *src, float *dst,
                                const QuadDev &q, const MeshDev &mesh,
                                const StageDev &stage, double time,
                                int dim_x, int collocation,
                                const double *efield, const DeformDev &geo,
                                hipStream_t stream)
{
  return dispatch<float>(D, ND, NQ, mode, src, dst, q, mesh, stage, time,
                         dim_x, collocation, efield, geo, stream);
}
} // namespace hd
