// hd_dirichlet.hip -- Dirichlet boundaries for the collapsed kernels: the
// boundary correction of a collapsed apply or LSRK stage (gfx950).
//
// The collapsed kernels read the inflow face of every cell from its
// neighbour, across the box edge as if the mesh were periodic.  At a
// Dirichlet boundary the reference (advection_operation.h:484-520, ECL;
// 1005-1100, FCL) uses the exterior value u+ = -u- + 2 g at the face
// quadrature points (u+ = -u- for homogeneous data).  With full upwinding an
// outflow face does not see u+, so only the inflow faces differ from the
// periodic operator, and the difference is linear: for a cell on the inflow
// boundary of direction d (side +1: inflow from the left, the cell's first
// layer; side -1: the last layer)
//   K(p) += b_d[p_d] ( -u(p_d = own inflow node, p_t) + 2 G(p_t)
//                      - u_wrap(p_d = its outflow node, p_t) )
// where u_wrap is the periodic image the kernel read, b_d the collapsed
// face column and G = (x)_{t != d} Sinv_t g(q_t): the face integral of g
// against the tangential basis, after the tangential inverse mass matrix
// (Sinv = (S^T W S)^-1 S^T W; S^-1 for n_points = k+1, I for collocation).
// One block per boundary cell of one direction (one to four waves, as many
// as the face has points; LDS for two copies of the face's quadrature
// values); one launch per direction, so that corner cells, which sit on
// several boundaries, are never updated by two blocks at once.  The
// correction reads only the input vector (not dst / ti_out / sol, which the
// kernel before it wrote), so it runs after the collapsed kernel:
// sol += b_dt C and ti_out += a_dt C in a stage.
#include "hd_device.h"
#include "hd_functions.h"

#include <algorithm>

namespace hd
{
namespace dirichlet
{
constexpr int kBlock = 256;
// face quadrature points per face (two LDS buffers of doubles)
constexpr int kMaxFacePoints = 7776;

__host__ __device__ inline long long ipow_ll(int b, int e)
{
  long long r = 1;
  for (int i = 0; i < e; ++i)
    r *= b;
  return r;
}

struct FaceDir
{
  int d;          // the direction
  int side;       // +1: inflow from the left (first layer), -1: last layer
  long long n_layer; // cells in the boundary layer
};

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void correct_kernel(
  const T *__restrict__ src, T *__restrict__ dst, const T *__restrict__ coef,
  MeshDev mesh, QuadDev q, StageDev stage, FaceDir fd, int D, int N, int NQ,
  int dim_x, double time)
{
  extern __shared__ double lds[];
  const int d = fd.d, side = fd.side;
  const int tid = threadIdx.x, nthr = blockDim.x;
  // the cell: layer index -> coordinates j != d, c_d on the boundary
  int c[kMaxDim];
  long long rem = blockIdx.x;
  long long cell = 0;
  for (int j = 0; j < D; ++j)
    {
      if (j == d)
        c[j] = side > 0 ? 0 : mesh.n[j] - 1;
      else
        {
          c[j] = (int)(rem % mesh.n[j]);
          rem /= mesh.n[j];
        }
      cell += (long long)c[j] * mesh.stride[j];
    }
  const long long wrap =
    cell + (side > 0 ? 1 : -1) * (long long)(mesh.n[d] - 1) * mesh.stride[d];
  const int nt = D - 1;
  const int face_nodes = (int)ipow_ll(N, nt);
  const long long cell_dofs = ipow_ll(N, D);
  const long long sd = ipow_ll(N, d); // node stride of direction d
  const int own_node = side > 0 ? 0 : N - 1;
  const int wrap_node = side > 0 ? N - 1 : 0;

  // G = (x)_t Sinv g(q_t), sum-factorised in LDS (inhomogeneous data)
  const double *G = nullptr;
  if (q.bc_inhomogeneous)
    {
      const int npts = (int)ipow_ll(NQ, nt);
      double *b0 = lds, *b1 = lds + npts;
      for (int i = tid; i < npts; i += nthr)
        {
          double x[kMaxDim];
          int r = i;
          for (int j = 0; j < D; ++j)
            {
              double local;
              if (j == d)
                local = side > 0 ? 0.0 : 1.0;
              else
                {
                  local = q.qp[r % NQ];
                  r /= NQ;
                }
              x[j] = mesh.lo[j] + (c[j] + local) * mesh.h[j];
            }
          b0[i] = eval_function(q.bc_function, q.bc_params, dim_x, D, x, time);
        }
      __syncthreads();
      // pass t contracts tangential index t (NQ -> N): layout
      // [a < N^t][i_t][b < NQ^(nt-1-t)] -> [a][p_t][b]
      double *in = b0, *out = b1;
      for (int t = 0; t < nt; ++t)
        {
          const int A = (int)ipow_ll(N, t);
          const int B = (int)ipow_ll(NQ, nt - 1 - t);
          const int n_out = A * N * B;
          for (int o = tid; o < n_out; o += nthr)
            {
              const int a = o % A, p = (o / A) % N, bb = o / (A * N);
              double s = 0.0;
              for (int i = 0; i < NQ; ++i)
                s += q.Sinv[p * NQ + i] * in[a + A * (i + NQ * bb)];
              out[o] = s;
            }
          __syncthreads();
          double *tmp = in;
          in = out;
          out = tmp;
        }
      G = in;
    }

  const int per = N * N + N;
  const T *bcol = coef + d * per + N * N;
  for (int f = tid; f < face_nodes; f += nthr)
    {
      // node index of face point f (tangential coordinates, first fastest)
      long long base = 0, m = 1;
      int r = f;
      for (int j = 0; j < D; ++j)
        {
          if (j != d)
            {
              base += (long long)(r % N) * m;
              r /= N;
            }
          m *= N;
        }
      const double own = (double)src[cell * cell_dofs + base + own_node * sd];
      const double wr = (double)src[wrap * cell_dofs + base + wrap_node * sd];
      const double diff = (G ? 2.0 * G[f] : 0.0) - own - wr;
      for (int pd = 0; pd < N; ++pd)
        {
          const long long idx = cell * cell_dofs + base + pd * sd;
          const double corr = (double)bcol[pd] * diff;
          if (MODE == 0)
            dst[idx] = (T)((double)dst[idx] + corr);
          else
            {
              stage.sol[idx] += stage.b_dt * corr;
              if (MODE == 1)
                {
                  T *ti = static_cast<T *>(stage.ti_out);
                  ti[idx] = (T)((double)ti[idx] + stage.a_dt * corr);
                }
            }
        }
    }
}

template <typename T>
hipError_t correct(int D, int N, int NQ, int mode, const T *src, T *dst,
                   const T *coef, const MeshDev &mesh, const InflowDev &inflow,
                   const QuadDev &q, const StageDev &stage, double time,
                   int dim_x, hipStream_t stream)
{
  const long long npts = dirichlet::ipow_ll(NQ, D - 1);
  const long long face_nodes = dirichlet::ipow_ll(N, D - 1);
  const size_t lds = q.bc_inhomogeneous ? 2 * npts * sizeof(double) : 0;
  const long long work = q.bc_inhomogeneous ? std::max(npts, face_nodes)
                                            : face_nodes;
  const int threads =
    (int)std::min<long long>(kBlock, (work + 63) / 64 * 64);
  for (int d = 0; d < D; ++d)
    {
      if (q.periodic[d] || inflow.side[d] == 0)
        continue;
      FaceDir fd;
      fd.d = d;
      fd.side = inflow.side[d];
      fd.n_layer = mesh.n_cells / mesh.n[d];
      const int m = mode == 1 && stage.ti_out ? 1 : (mode == 0 ? 0 : 2);
#define HD_DIR_LAUNCH(MM)                                                      \
  {                                                                            \
    hipError_t e = hipFuncSetAttribute(                                        \
      reinterpret_cast<const void *>(&correct_kernel<T, MM>),                  \
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                   \
    if (e != hipSuccess)                                                       \
      return e;                                                                \
    hipLaunchKernelGGL((correct_kernel<T, MM>), dim3((unsigned)fd.n_layer),    \
                       dim3(threads), lds, stream, src, dst, coef, mesh, q,    \
                       stage, fd, D, N, NQ, dim_x, time);                      \
  }
      if (m == 0)
        HD_DIR_LAUNCH(0)
      else if (m == 1)
        HD_DIR_LAUNCH(1)
      else
        HD_DIR_LAUNCH(2)
#undef HD_DIR_LAUNCH
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess)
        return e;
    }
  return hipSuccess;
}
} // namespace dirichlet

bool dirichlet_collapsed_supported(int D, int NQ, bool inhomogeneous)
{
  return D >= 2 && D <= kMaxDim && NQ <= 8 &&
         (!inhomogeneous ||
          dirichlet::ipow_ll(NQ, D - 1) <= dirichlet::kMaxFacePoints);
}

hipError_t dirichlet_correct_f64(int D, int N, int NQ, int mode,
                                 const double *src, double *dst,
                                 const double *coef, const MeshDev &mesh,
                                 const InflowDev &inflow, const QuadDev &q,
                                 const StageDev &stage, double time,
                                 int dim_x, hipStream_t stream)
{
  return dirichlet::correct<double>(D, N, NQ, mode, src, dst, coef, mesh,
                                    inflow, q, stage, time, dim_x, stream);
}

hipError_t dirichlet_correct_f32(int D, int N, int NQ, int mode,
                                 const float *src, float *dst,
                                 const float *coef, const MeshDev &mesh,
                                 const InflowDev &inflow, const QuadDev &q,
                                 const StageDev &stage, double time,
                                 int dim_x, hipStream_t stream)
{
  return dirichlet::correct<float>(D, N, NQ, mode, src, dst, coef, mesh,
                                   inflow, q, stage, time, dim_x, stream);
}
} // namespace hd
