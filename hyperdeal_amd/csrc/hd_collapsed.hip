// hd_collapsed.hip -- constant-velocity DG advection apply on periodic
// Cartesian phase space, gfx950.
//
// For a constant velocity on an affine Cartesian mesh with N >= k+1 Gauss
// points (or Gauss-Lobatto collocation) every integral of the reference's ECL
// operator (operators/advection/advection_operation.h:221-566) factorises
// over the directions, and the inverse mass matrix cancels the tangential
// factors exactly.  The apply is therefore
//
//   dst(p) = sum_d [ sum_j A_d[p_d][j] src(p with p_d = j)
//                    + b_d[p_d] src_nb,d(p with p_d = inflow node) ]
//
// with A_d (N x N) and b_d (N) the 1D operator of direction d (built on the
// host by hd::collapse_direction) and src_nb,d the upwind neighbour across d.
// The outflow neighbour never enters the upwind flux.
//
// Work decomposition (one cell of N^D values per block for 3+3, several for
// smaller D): a thread owns an N x N block of points in a pair of directions
// (layout p owns directions 2p, 2p+1), so the 1D contractions of both run in
// registers; the cell is staged in LDS (coalesced global loads, padded by one
// value per N^2 to keep the strided layout reads bank-conflict free) and the
// per-layout contributions are summed through the same LDS buffer.
#include "hd_device.h"

namespace hd
{
namespace collapsed
{
constexpr int kBlock = 256;

template <int D, int N>
struct Shape
{
  static constexpr int NN = N * N;
  static constexpr int CELL = ipow(N, D);
  static constexpr int TPC = ipow(N, D - 2);   // threads per cell
  static constexpr int CPB = kBlock / TPC;      // cells per block
  static constexpr int THREADS = CPB * TPC;
  static constexpr int CELLP = CELL + CELL / NN; // padded cell in LDS
  static constexpr int P = (D + 1) / 2;          // layouts
  static_assert(TPC <= kBlock, "cell too large for one block");
};

__device__ constexpr int lay_a(int p) { return 2 * p; }
template <int D>
__device__ constexpr int lay_b(int p)
{
  return 2 * p + 1 < D ? 2 * p + 1 : (2 * p == 0 ? 1 : 0);
}
template <int D>
__device__ constexpr bool lay_b_active(int p)
{
  return 2 * p + 1 < D;
}

// 16-byte (or smaller, when a cell is not a multiple of 16 bytes) chunk
template <typename T, int VEC>
struct alignas(VEC * sizeof(T)) Chunk
{
  T v[VEC];
};

template <typename T, int CELL>
__host__ __device__ constexpr int chunk_width()
{
  return (CELL % (16 / (int)sizeof(T)) == 0) ? 16 / (int)sizeof(T) : 1;
}

template <int N>
__device__ constexpr int padded(int i)
{
  return i + i / (N * N);
}

// index of the thread's (i_a, i_b) = (0, 0) point for layout (a, b)
template <int D, int N>
__device__ inline int layout_base(int tau, int a, int b)
{
  int base = 0, pw = 1;
#pragma unroll
  for (int j = 0; j < D; ++j)
    {
      if (j != a && j != b)
        {
          base += (tau % N) * pw;
          tau /= N;
        }
      pw *= N;
    }
  return base;
}

// Contributions of directions (a [, b]) for one thread's N x N block.
//   u: the block's values, r = i_a + N i_b; out: accumulated contributions
template <int D, int N, typename T>
__device__ inline void block_contrib(int p, int base, long long cell,
                                     const int *coord,
                                     const T *__restrict__ src,
                                     const T *__restrict__ coef,
                                     const MeshDev &mesh,
                                     const InflowDev &inflow, const T *u,
                                     T *out)
{
  constexpr int NN = N * N;
  constexpr int CELL = ipow(N, D);
  const int a = lay_a(p), b = lay_b<D>(p);
  const int sa = ipow(N, a), sb = ipow(N, b);
#pragma unroll
  for (int pass = 0; pass < 2; ++pass)
    {
      if (pass == 1 && !lay_b_active<D>(p))
        break;
      const int d = pass == 0 ? a : b;
      const int side = inflow.side[d];
      if (side == 0)
        continue;
      const T *A = coef + d * (NN + N);
      const T *bv = A + NN;
      // inflow face of the upwind neighbour: its node p_d = N-1 (from the
      // left) or 0 (from the right), other indices equal
      const long long st = mesh.stride[d];
      const int nd = mesh.n[d];
      const long long nb =
        side > 0 ? (coord[d] == 0 ? cell + (nd - 1) * st : cell - st)
                 : (coord[d] == nd - 1 ? cell - (nd - 1) * st : cell + st);
      const int fnode = side > 0 ? N - 1 : 0;
      const T *nsrc = src + nb * CELL + base + fnode * (pass == 0 ? sa : sb);
#pragma unroll
      for (int l = 0; l < N; ++l) // line index: i_b (pass 0) or i_a (pass 1)
        {
          const T face = nsrc[l * (pass == 0 ? sb : sa)];
          T line[N];
#pragma unroll
          for (int j = 0; j < N; ++j)
            line[j] = pass == 0 ? u[j + N * l] : u[l + N * j];
#pragma unroll
          for (int i = 0; i < N; ++i)
            {
              T acc = bv[i] * face;
#pragma unroll
              for (int j = 0; j < N; ++j)
                acc += A[i * N + j] * line[j];
              if (pass == 0)
                out[i + N * l] += acc;
              else
                out[l + N * i] += acc;
            }
        }
    }
}

// MODE 0: dst = apply(src).  MODE 1: LSRK stage with K = apply(src).
template <int D, int N, typename T, int MODE>
__global__ __launch_bounds__(kBlock) void apply_kernel(
  const T *__restrict__ src, T *__restrict__ dst, const T *__restrict__ coef,
  MeshDev mesh, InflowDev inflow, StageDev stage, long long n_logical)
{
  using S = Shape<D, N>;
  constexpr int NN = S::NN, CELL = S::CELL, TPC = S::TPC, CPB = S::CPB;
  constexpr int CELLP = S::CELLP, P = S::P;
  __shared__ T lds[CPB * CELLP];

  const long long lb = xcd_chunked_block(n_logical);
  if (lb >= n_logical)
    return; // whole block idle: uniform exit before any barrier
  const int tid = threadIdx.x;
  const long long cell0 = lb * CPB;
  const int ncell = (int)min((long long)CPB, mesh.n_cells - cell0);

  // 1) coalesced load of the block's cells into LDS (16-byte chunks)
  constexpr int VEC = chunk_width<T, CELL>();
  using V = Chunk<T, VEC>;
  {
    const V *g = reinterpret_cast<const V *>(src + cell0 * CELL);
    const int nchunks = ncell * CELL / VEC;
    for (int c = tid; c < nchunks; c += S::THREADS)
      {
        const V v = g[c];
        const int i = c * VEC; // block-local value index
        const int lc = i / CELL, li = i - lc * CELL;
        T *dstl = lds + lc * CELLP + padded<N>(li);
#pragma unroll
        for (int e = 0; e < VEC; ++e)
          dstl[e] = v.v[e];
      }
  }
  __syncthreads();

  const int lc = tid / TPC, tau = tid - lc * TPC;
  const bool active = tid < S::THREADS && lc < ncell;
  const long long cell = CPB == 1 ? cell0 : cell0 + lc;
  T *cl = lds + lc * CELLP;
  // cell coordinates (cell ids fit in 32 bits; checked on the host)
  int coord[D];
  {
    unsigned rem = (unsigned)cell;
#pragma unroll
    for (int d = 0; d < D; ++d)
      {
        const unsigned q = rem / (unsigned)mesh.n[d];
        coord[d] = (int)(rem - q * (unsigned)mesh.n[d]);
        rem = q;
      }
  }

  // 2) every layout's values into registers
  T u[P][NN];
  int base[P];
#pragma unroll
  for (int p = 0; p < P; ++p)
    {
      const int a = lay_a(p), b = lay_b<D>(p);
      base[p] = layout_base<D, N>(tau, a, b);
      const int sa = ipow(N, a), sb = ipow(N, b);
#pragma unroll
      for (int r = 0; r < NN; ++r)
        u[p][r] = active ? cl[padded<N>(base[p] + (r % N) * sa + (r / N) * sb)]
                         : T(0);
    }
  __syncthreads();

  // 3) contributions per layout, summed in LDS
#pragma unroll
  for (int p = 0; p < P; ++p)
    {
      T out[NN];
#pragma unroll
      for (int r = 0; r < NN; ++r)
        out[r] = T(0);
      if (active)
        block_contrib<D, N, T>(p, base[p], cell, coord, src, coef, mesh,
                               inflow, u[p], out);
      const int a = lay_a(p), b = lay_b<D>(p);
      const int sa = ipow(N, a), sb = ipow(N, b);
      if (active)
#pragma unroll
        for (int r = 0; r < NN; ++r)
          {
            T *q = cl + padded<N>(base[p] + (r % N) * sa + (r / N) * sb);
            if (p == 0)
              *q = out[r];
            else
              *q += out[r];
          }
      __syncthreads();
    }

  // 4) coalesced epilogue
  {
    const int nchunks = ncell * CELL / VEC;
    for (int c = tid; c < nchunks; c += S::THREADS)
      {
        const int i = c * VEC;
        const int lcc = i / CELL, li = i - lcc * CELL;
        const T *k = lds + lcc * CELLP + padded<N>(li);
        const long long g = cell0 * CELL + i;
        if (MODE == 0)
          {
            V v;
#pragma unroll
            for (int e = 0; e < VEC; ++e)
              v.v[e] = k[e];
            reinterpret_cast<V *>(dst)[g / VEC] = v;
          }
        else
          {
            double s[VEC];
#pragma unroll
            for (int e = 0; e < VEC; ++e)
              s[e] = stage.sol[g + e];
#pragma unroll
            for (int e = 0; e < VEC; ++e)
              stage.sol[g + e] = s[e] + stage.b_dt * (double)k[e];
            if (stage.ti_out)
              {
                V v;
#pragma unroll
                for (int e = 0; e < VEC; ++e)
                  v.v[e] = (T)(s[e] + stage.a_dt * (double)k[e]);
                reinterpret_cast<V *>(stage.ti_out)[g / VEC] = v;
              }
          }
      }
  }
}

template <int D, int N, typename T>
hipError_t launch(int mode, const T *src, T *dst, const T *coef,
                  const MeshDev &mesh, const InflowDev &inflow,
                  const StageDev &stage, hipStream_t stream)
{
  using S = Shape<D, N>;
  const long long n_logical = (mesh.n_cells + S::CPB - 1) / S::CPB;
  const long long grid = 8 * ((n_logical + 7) / 8);
  if (mode == 0)
    hipLaunchKernelGGL((apply_kernel<D, N, T, 0>), dim3((unsigned)grid),
                       dim3(kBlock), 0, stream, src, dst, coef, mesh, inflow,
                       stage, n_logical);
  else
    hipLaunchKernelGGL((apply_kernel<D, N, T, 1>), dim3((unsigned)grid),
                       dim3(kBlock), 0, stream, src, dst, coef, mesh, inflow,
                       stage, n_logical);
  return hipGetLastError();
}

template <typename T>
hipError_t dispatch(int D, int N, int mode, const T *src, T *dst,
                    const T *coef, const MeshDev &mesh,
                    const InflowDev &inflow, const StageDev &stage,
                    hipStream_t stream)
{
#define HD_CASE(DD, NN_)                                                     \
  if (D == DD && N == NN_)                                                   \
    return launch<DD, NN_, T>(mode, src, dst, coef, mesh, inflow, stage,     \
                              stream);
  HD_CASE(2, 2) HD_CASE(2, 3) HD_CASE(2, 4) HD_CASE(2, 5) HD_CASE(2, 6)
  HD_CASE(2, 7) HD_CASE(2, 8)
  HD_CASE(3, 2) HD_CASE(3, 3) HD_CASE(3, 4) HD_CASE(3, 5) HD_CASE(3, 6)
  HD_CASE(4, 2) HD_CASE(4, 3) HD_CASE(4, 4) HD_CASE(4, 5) HD_CASE(4, 6)
  HD_CASE(5, 2) HD_CASE(5, 3) HD_CASE(5, 4)
  HD_CASE(6, 2) HD_CASE(6, 3) HD_CASE(6, 4)
#undef HD_CASE
  return hipErrorInvalidValue;
}

} // namespace collapsed

bool collapsed_supported(int D, int N)
{
  if (D < 2 || D > 6 || N < 2)
    return false;
  static const int maxN[7] = {0, 0, 8, 6, 6, 4, 4};
  return N <= maxN[D];
}

hipError_t collapsed_apply_f64(int D, int N, int mode, const double *src,
                               double *dst, const double *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream)
{
  return collapsed::dispatch<double>(D, N, mode, src, dst, coef, mesh, inflow,
                                     stage, stream);
}

hipError_t collapsed_apply_f32(int D, int N, int mode, const float *src,
                               float *dst, const float *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream)
{
  return collapsed::dispatch<float>(D, N, mode, src, dst, coef, mesh, inflow,
                                    stage, stream);
}
} // namespace hd
