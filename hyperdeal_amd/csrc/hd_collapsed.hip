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
// Work decomposition.  A block of 256 threads works on CPB cells at a time
// (one cell of N^D values for 3+3, k=3).  A thread owns an N x N block of
// points in a pair of directions (layout p owns directions 2p, 2p+1), so the
// 1D contractions of both run in registers.  The cell sits in LDS (U, padded
// by one value per N^2 so the three strided layouts read without bank
// conflicts); each layout adds its contribution into a second LDS buffer R,
// which is stored with coalesced 16-byte accesses.
//
// Blocks are persistent (two per CU) and walk an XCD-local range of cells:
// while a block computes one cell, the next cell's values and all its inflow
// face values are already in flight into registers, so HBM and L2 latency
// overlap with the arithmetic.  The logical cells of one XCD are handed out
// round-robin to its blocks, so the blocks of an XCD work on neighbouring
// cells at the same time and face reads hit its L2 or the MALL.
#include "hd_device.h"

#include <type_traits>

namespace hd
{
namespace collapsed
{
constexpr int kBlock = 256;

// LDS per block: two padded buffers of CPB cells, at most this many bytes so
// that two blocks fit on a CU (160 KB)
constexpr int kLdsPerBlock = 76 * 1024;

template <int D, int N, typename T>
struct Shape
{
  static constexpr int NN = N * N;
  static constexpr int CELL = ipow(N, D);
  static constexpr int TPC = ipow(N, D - 2);    // threads per cell
  static constexpr int CELLP = CELL + CELL / NN; // padded cell in LDS
  static constexpr int CPB_THREADS = kBlock / TPC;
  static constexpr int CPB_LDS =
    kLdsPerBlock / (2 * CELLP * (int)sizeof(T)) > 0
      ? kLdsPerBlock / (2 * CELLP * (int)sizeof(T))
      : 1;
  static constexpr int CPB = CPB_THREADS < CPB_LDS ? CPB_THREADS : CPB_LDS;
  static constexpr int THREADS = CPB * TPC;       // active threads
  static constexpr int P = (D + 1) / 2;           // layouts
  static_assert(TPC <= kBlock, "cell too large for one block");
};

__host__ __device__ constexpr int lay_a(int p) { return 2 * p; }
template <int D>
__host__ __device__ constexpr int lay_b(int p)
{
  return 2 * p + 1 < D ? 2 * p + 1 : (2 * p == 0 ? 1 : 0);
}
template <int D>
__host__ __device__ constexpr bool lay_b_active(int p)
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

// Registers a thread carries from one cell to the next: its share of the
// cell's values (coalesced chunks) and the inflow face values of its lines.
template <int D, int N, typename T>
struct Prefetch
{
  using S = Shape<D, N, T>;
  static constexpr int VEC = chunk_width<T, S::CELL>();
  static constexpr int CHUNKS = S::CPB * S::CELL / VEC;
  static constexpr int CPT = (CHUNKS + kBlock - 1) / kBlock;
  Chunk<T, VEC> c[CPT];
  T face[S::P][2][N];
};

template <int D, int N, typename T>
__device__ inline void issue_chunks(Prefetch<D, N, T> &pf, const T *src,
                                    long long cell0, int ncell)
{
  using S = Shape<D, N, T>;
  using PF = Prefetch<D, N, T>;
  constexpr int CELL = S::CELL, VEC = PF::VEC;
  const Chunk<T, VEC> *g =
    reinterpret_cast<const Chunk<T, VEC> *>(src + cell0 * CELL);
  const int nchunks = ncell * CELL / VEC;
#pragma unroll
  for (int k = 0; k < PF::CPT; ++k)
    {
      const int c = threadIdx.x + k * kBlock;
      if (c < nchunks)
        pf.c[k] = g[c];
    }
}

// cell coordinates (cell ids fit in 32 bits; checked on the host)
template <int D>
__device__ inline void cell_coords(long long cell, const MeshDev &mesh,
                                   int *coord)
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

// the inflow face values of layout p's lines of `cell`
template <int D, int N, typename T>
__device__ inline void issue_faces(Prefetch<D, N, T> &pf, int p,
                                   const T *src, long long cell,
                                   const int *coord, int base,
                                   const MeshDev &mesh,
                                   const InflowDev &inflow)
{
  constexpr int CELL = ipow(N, D);
#pragma unroll
  for (int pass = 0; pass < 2; ++pass)
    {
      if (pass == 1 && !lay_b_active<D>(p))
        break;
      const int a = lay_a(p), b = lay_b<D>(p);
      const int d = pass == 0 ? a : b;
      const int side = inflow.side[d];
      if (side == 0)
        continue;
      const int sa = ipow(N, a), sb = ipow(N, b);
      const long long st = mesh.stride[d];
      const int nd = mesh.n[d];
      const long long nb =
        side > 0 ? (coord[d] == 0 ? cell + (nd - 1) * st : cell - st)
                 : (coord[d] == nd - 1 ? cell - (nd - 1) * st : cell + st);
      const int fnode = side > 0 ? N - 1 : 0;
      const T *ns = src + nb * CELL + base + fnode * (pass == 0 ? sa : sb);
#pragma unroll
      for (int l = 0; l < N; ++l)
        pf.face[p][pass][l] = ns[l * (pass == 0 ? sb : sa)];
    }
}

// Contributions of directions (a [, b]) for one thread's N x N block.
//   u: the block's values, r = i_a + N i_b; out: contributions
template <int D, int N, typename T>
__device__ inline void block_contrib(int p, const T *__restrict__ coef,
                                     const InflowDev &inflow, const T *u,
                                     const T (*face)[N], T *out)
{
  constexpr int NN = N * N;
  const int a = lay_a(p), b = lay_b<D>(p);
#pragma unroll
  for (int pass = 0; pass < 2; ++pass)
    {
      if (pass == 1 && !lay_b_active<D>(p))
        break;
      const int d = pass == 0 ? a : b;
      if (inflow.side[d] == 0)
        continue;
      const T *A = coef + d * (NN + N);
      const T *bv = A + NN;
#pragma unroll
      for (int l = 0; l < N; ++l) // line index: i_b (pass 0) or i_a (pass 1)
        {
          const T f = face[pass][l];
#pragma unroll
          for (int i = 0; i < N; ++i)
            {
              T acc = bv[i] * f;
#pragma unroll
              for (int j = 0; j < N; ++j)
                acc += A[i * N + j] * (pass == 0 ? u[j + N * l] : u[l + N * j]);
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
__global__ __launch_bounds__(kBlock, 2) void apply_kernel(
  const T *__restrict__ src, T *__restrict__ dst, const T *__restrict__ coef,
  MeshDev mesh, InflowDev inflow, StageDev stage, long long n_logical,
  int slots)
{
  using S = Shape<D, N, T>;
  using PF = Prefetch<D, N, T>;
  constexpr int NN = S::NN, CELL = S::CELL, TPC = S::TPC, CPB = S::CPB;
  constexpr int CELLP = S::CELLP, P = S::P, VEC = PF::VEC;
  __shared__ T U[CPB * CELLP];
  __shared__ T R[CPB * CELLP];

  // this block's share of its XCD's contiguous range of logical cells
  const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const long long per = (n_logical + 7) / 8;
  const long long lend = min(n_logical, (xcd + 1) * per);
  long long lb = xcd * per + slot;
  if (lb >= lend)
    return; // uniform for the block, before any barrier

  const int tid = threadIdx.x;
  const int lc = tid / TPC, tau = tid - lc * TPC;
  const int lbase = lc * CELLP;
  int base[P];
#pragma unroll
  for (int p = 0; p < P; ++p)
    base[p] = layout_base<D, N>(tau, lay_a(p), lay_b<D>(p));

  PF pf;
  {
    const long long c0 = lb * CPB;
    const int nc = (int)min((long long)CPB, mesh.n_cells - c0);
    issue_chunks<D, N, T>(pf, src, c0, nc);
    if (tid < S::THREADS && lc < nc)
      {
        const long long cell = CPB == 1 ? c0 : c0 + lc;
        int coord[D];
        cell_coords<D>(cell, mesh, coord);
#pragma unroll
        for (int p = 0; p < P; ++p)
          issue_faces<D, N, T>(pf, p, src, cell, coord, base[p], mesh,
                               inflow);
      }
  }

  for (; lb < lend; lb += slots)
    {
      const long long cell0 = lb * CPB;
      const int ncell = (int)min((long long)CPB, mesh.n_cells - cell0);
      const bool active = tid < S::THREADS && lc < ncell;
      const int nchunks = ncell * CELL / VEC;
      // 1) this cell's values into LDS, keep its faces
#pragma unroll
      for (int k = 0; k < PF::CPT; ++k)
        {
          const int c = tid + k * kBlock;
          if (c < nchunks)
            {
              const int i = c * VEC;
              const int l = i / CELL, li = i - l * CELL;
              T *dl = U + l * CELLP + padded<N>(li);
#pragma unroll
              for (int e = 0; e < VEC; ++e)
                dl[e] = pf.c[k].v[e];
            }
        }
      __syncthreads();

      // 2) the next cell's values go out now and land during the arithmetic;
      // its face values follow layout by layout, as soon as the current
      // layout has consumed its own
      const long long ln = lb + slots;
      const bool next = ln < lend;
      long long ncell0 = 0, nxt_cell = 0;
      bool nxt_active = false;
      int ncoord[D];
      if (next)
        {
          ncell0 = ln * CPB;
          const int nc = (int)min((long long)CPB, mesh.n_cells - ncell0);
          issue_chunks<D, N, T>(pf, src, ncell0, nc);
          nxt_active = tid < S::THREADS && lc < nc;
          nxt_cell = CPB == 1 ? ncell0 : ncell0 + lc;
          if (nxt_active)
            cell_coords<D>(nxt_cell, mesh, ncoord);
        }

      // 3) per layout: read the block from U, add its directions into R
#pragma unroll
      for (int p = 0; p < P; ++p)
        {
          const int sa = ipow(N, lay_a(p)), sb = ipow(N, lay_b<D>(p));
          if (active)
            {
              T u[NN], out[NN];
#pragma unroll
              for (int r = 0; r < NN; ++r)
                {
                  u[r] = U[lbase +
                           padded<N>(base[p] + (r % N) * sa + (r / N) * sb)];
                  out[r] = T(0);
                }
              block_contrib<D, N, T>(p, coef, inflow, u, pf.face[p], out);
#pragma unroll
              for (int r = 0; r < NN; ++r)
                {
                  T *q = R + lbase +
                         padded<N>(base[p] + (r % N) * sa + (r / N) * sb);
                  if (p == 0)
                    *q = out[r];
                  else
                    *q += out[r];
                }
            }
          if (nxt_active)
            issue_faces<D, N, T>(pf, p, src, nxt_cell, ncoord, base[p], mesh,
                                 inflow);
          __syncthreads();
        }

      // 4) coalesced epilogue from R.  U is rewritten at the top of the next
      // iteration only after every thread passed the layout barriers, and R
      // only in the next iteration's first layout, after the barrier that
      // follows the U write.
#pragma unroll
      for (int k = 0; k < PF::CPT; ++k)
        {
          const int c = tid + k * kBlock;
          if (c >= nchunks)
            continue;
          const int i = c * VEC;
          const int l = i / CELL, li = i - l * CELL;
          const T *kv = R + l * CELLP + padded<N>(li);
          const long long g = cell0 * CELL + i;
          if (MODE == 0)
            {
              Chunk<T, VEC> v;
#pragma unroll
              for (int e = 0; e < VEC; ++e)
                v.v[e] = kv[e];
              reinterpret_cast<Chunk<T, VEC> *>(dst)[g / VEC] = v;
            }
          else
            {
              double s[VEC];
#pragma unroll
              for (int e = 0; e < VEC; ++e)
                s[e] = stage.sol[g + e];
#pragma unroll
              for (int e = 0; e < VEC; ++e)
                stage.sol[g + e] = s[e] + stage.b_dt * (double)kv[e];
              if (stage.ti_out)
                {
                  Chunk<T, VEC> v;
#pragma unroll
                  for (int e = 0; e < VEC; ++e)
                    v.v[e] = (T)(s[e] + stage.a_dt * (double)kv[e]);
                  reinterpret_cast<Chunk<T, VEC> *>(stage.ti_out)[g / VEC] = v;
                }
            }
        }
    }
}

int g_slots_per_xcd = 0;

int slots_per_xcd()
{
  if (g_slots_per_xcd == 0)
    {
      int dev = 0, cus = 256;
      if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus,
                                    hipDeviceAttributeMultiprocessorCount, dev);
      g_slots_per_xcd = 2 * ((cus + 7) / 8); // two blocks per CU
    }
  return g_slots_per_xcd;
}

template <int D, int N, typename T>
hipError_t launch(int mode, const T *src, T *dst, const T *coef,
                  const MeshDev &mesh, const InflowDev &inflow,
                  const StageDev &stage, hipStream_t stream)
{
  using S = Shape<D, N, T>;
  const long long n_logical = (mesh.n_cells + S::CPB - 1) / S::CPB;
  // no more blocks per XCD than the XCD has logical cells
  const long long per = (n_logical + 7) / 8;
  const int used = (int)min((long long)slots_per_xcd(), per);
  const unsigned grid = 8u * (unsigned)used;
  if (mode == 0)
    hipLaunchKernelGGL((apply_kernel<D, N, T, 0>), dim3(grid), dim3(kBlock),
                       0, stream, src, dst, coef, mesh, inflow, stage,
                       n_logical, used);
  else
    hipLaunchKernelGGL((apply_kernel<D, N, T, 1>), dim3(grid), dim3(kBlock),
                       0, stream, src, dst, coef, mesh, inflow, stage,
                       n_logical, used);
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
