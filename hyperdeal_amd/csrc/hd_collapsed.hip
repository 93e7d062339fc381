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
// (one cell of N^D values for 3+3, k=3).  A thread owns an N x N tile of
// points in a pair of directions (a "layout"), so the 1D contractions of both
// directions run in registers; the D directions are covered by (D+1)/2
// layouts.  The cell sits in LDS (U) in a layout whose strides are chosen so
// that every layout reads it without bank conflicts; the layouts
// but the last add their contributions into a second LDS buffer R, and the
// last layout adds R to its own contribution and stores to HBM directly (its
// threads own points that are contiguous in memory).
//
// Latency.  Two (FP64) or four (FP32) blocks are resident per CU and each
// walks several cells: for 3+3, k=3 one row of cells along x0 in upwind
// order (the next cell's x0 inflow face is then the current cell's outflow
// face, already in LDS), one block per row so that the dispatcher hands the
// rows of an XCD out in order; for other shapes persistent blocks over an
// XCD-local range.  While a block computes one cell, the next cell's values
// and inflow faces are already in flight into registers: the loads are
// unconditional (the last iteration re-reads its own cell) and the barriers
// order LDS only, so the compiler waits for each load with a counted vmcnt
// instead of draining the queue.  The blocks of an XCD work on neighbouring
// cells at the same time, so face reads hit its L2 or the MALL.
#include "hd_device.h"

#include <algorithm>
#include <type_traits>
#include <utility>

namespace hd
{
namespace collapsed
{
constexpr int kBlock = 256;

// Orders the block's LDS accesses.  __syncthreads() would also wait for every
// outstanding global load, i.e. for the prefetch of the next cell.
__device__ inline void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int I>
using ic = std::integral_constant<int, I>;

template <typename F, int... I>
__device__ inline void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
  (f(ic<I>{}), ...);
}

// f(ic<0>), ..., f(ic<M-1>)
template <int M, typename F>
__device__ inline void static_for(F &&f)
{
  static_for_impl(f, std::make_integer_sequence<int, M>{});
}

// Layouts and the LDS layout of a cell.  Layout p owns directions a(p) and
// b(p) (b only as a tile index when !b_active(p)).  Element (i_0..i_{D-1})
// of a cell sits at LDS offset sum_d i_d W(d).
//
// 3+3, N=4: layouts (0,4), (1,5), (2,3), unpadded (W = 4^d).  gfx950's LDS
// (64 banks of 4 bytes) checks a ds_read/write_b32 for bank conflicts in
// groups of 16 lanes and a b64 / b128 access in groups of 8 lanes (measured);
// under these rules every single-element tile access of the three layouts,
// the 16-byte row accesses of the (0,4) layout along direction 0 and the
// 16-byte chunk writes are conflict-free.  The paired forms
// (ds_read2/write2_b64) check 16 lanes per address slot and would conflict,
// so single-element tile accesses are volatile, which keeps the compiler from
// pairing them.  The (0,4) layout's inflow faces of direction 4 are
// contiguous in memory, and the (2,3) layout's points are contiguous runs of
// 16 values, so its stores coalesce.
// Other shapes: layouts (0,1), (2,3), ..., W = N^d + N^(d-2) for d >= 2.
template <int D, int N>
struct Plan
{
  static constexpr bool kTuned = (D == 6 && N == 4);
  static constexpr int P = (D + 1) / 2;
  __host__ __device__ static constexpr int a(int p)
  {
    return kTuned ? p : 2 * p;
  }
  __host__ __device__ static constexpr int b(int p)
  {
    return kTuned ? (p == 0 ? 4 : p == 1 ? 5 : 3)
                  : (2 * p + 1 < D ? 2 * p + 1 : (p == 0 ? 1 : 0));
  }
  __host__ __device__ static constexpr bool b_active(int p)
  {
    return kTuned || 2 * p + 1 < D;
  }
  __host__ __device__ static constexpr int W(int d)
  {
    return ipow(N, d) + (!kTuned && d >= 2 ? ipow(N, d - 2) : 0);
  }
  __host__ __device__ static constexpr int span()
  {
    int s = 0;
    for (int d = 0; d < D; ++d)
      s += (N - 1) * W(d);
    return s + 1;
  }
  // padded cell size, a multiple of N so that cells keep row alignment
  static constexpr int CELLP = (span() + N - 1) / N * N;
  // every W(d), d >= 1, is a multiple of N: rows along direction 0 start at
  // multiples of N
  static constexpr bool kAlignedRows = kTuned;
  // Row walk: a block owns G cells that are neighbours along direction 1
  // (one per 256 threads) and walks their rows along direction 0 in upwind
  // order.  The direction-0 inflow face of the next cell is then the current
  // cell's outflow face, and the direction-1 inflow face of all but one of
  // the G cells is the outflow face of a cell of the same block: both are
  // read from LDS instead of HBM / L2.
  static constexpr bool kRowWalk = kTuned;
};

// G: cells per block along direction 1 (row walk only).  BT: threads per
// block; 0 trims the block to the waves that own points (a cell group that
// LDS limits to fewer than kBlock threads then runs no shadow waves and its
// launch bounds leave the tiles more registers); kBlock for kernels whose
// block size is fixed (batches).  LB: blocks per CU that share the LDS (0:
// two for FP64, four for FP32).
template <int D, int N, typename T, int G = 1, int BT = 0, int LB = 0>
struct Shape
{
  using PL = Plan<D, N>;
  static constexpr int P = PL::P;
  static constexpr int NN = N * N;
  static constexpr int CELL = ipow(N, D);
  static constexpr int TPC = ipow(N, D - 2);   // threads per cell
  static constexpr int CELLP = PL::CELLP;
  static constexpr bool kGroup = PL::kRowWalk;
  static_assert(kGroup || G == 1, "cell groups need the row walk");
  // LDS budget per block: that of the most resident blocks per CU
  static constexpr int kLds =
    160 * 1024 / (LB > 0 ? LB : (sizeof(T) == 8 ? 2 : 4)) - 2048;
  static constexpr int NBUF = P > 1 ? 2 : 1;
  static constexpr int CPB_THREADS = (BT > 0 ? BT : kBlock) / TPC;
  static constexpr int CPB_LDS =
    kLds / (NBUF * CELLP * (int)sizeof(T)) > 0
      ? kLds / (NBUF * CELLP * (int)sizeof(T))
      : 1;
  static constexpr int CPB =
    kGroup ? G : (CPB_THREADS < CPB_LDS ? CPB_THREADS : CPB_LDS);
  static constexpr int THREADS = CPB * TPC; // threads that own points
  static constexpr int BLOCK =
    kGroup ? G * TPC : (BT > 0 ? BT : (THREADS + 63) / 64 * 64);
  static constexpr bool kAllOwn = THREADS == BLOCK;
  // 16-byte chunks of the coalesced cell loads: per cell (groups) or over
  // the block's consecutive cells
  static constexpr int VEC =
    CELL % (16 / (int)sizeof(T)) == 0 ? 16 / (int)sizeof(T) : 1;
  static constexpr int CHUNKS = CPB * CELL / VEC;
  static constexpr int CPT =
    kGroup ? CELL / VEC / TPC : (CHUNKS + BLOCK - 1) / BLOCK;
  // a chunk is one aligned run in LDS
  static constexpr bool kChunkRuns = PL::kAlignedRows && N % VEC == 0;
  // rows along direction 0 as 16-byte LDS / global vectors
  static constexpr int RV = 16 / (int)sizeof(T);
  static constexpr bool kRowVec = PL::kAlignedRows && N % RV == 0;
  static constexpr bool kFaceVec = N % RV == 0;
  static_assert(TPC <= kBlock, "cell too large for one block");
  static_assert(THREADS <= BLOCK, "block smaller than its cells");
  static_assert(!kGroup || CELL % (VEC * TPC) == 0, "chunks per cell");
};

template <typename T, int V>
struct alignas(V * sizeof(T)) Vec
{
  T v[V];
};

// LDS offset of the in-cell lexicographic index i
template <int D, int N>
__device__ inline int lds_offset(int i)
{
  int off = 0;
#pragma unroll
  for (int d = 0; d < D; ++d)
    {
      off += (i % N) * Plan<D, N>::W(d);
      i /= N;
    }
  return off;
}

// global (in-cell) and LDS offsets of the (0, 0) point of the tile that
// thread tau owns in the layout with tile directions (a, b)
template <int D, int N>
__device__ inline void tile_base(int tau, int a, int b, int &g, int &l)
{
  g = 0;
  l = 0;
  int pw = 1;
#pragma unroll
  for (int d = 0; d < D; ++d)
    {
      if (d != a && d != b)
        {
          const int dig = tau % N;
          tau /= N;
          g += dig * pw;
          l += dig * Plan<D, N>::W(d);
        }
      pw *= N;
    }
}

// tile element r = i_a + N i_b of layout p: read from an LDS buffer
template <int D, int N, typename T, int p>
__device__ inline void tile_read(const T *buf, int base, T *u)
{
  using S = Shape<D, N, T>;
  using PL = Plan<D, N>;
  constexpr int a = PL::a(p), b = PL::b(p);
  if constexpr (a == 0 && S::kRowVec)
    {
      constexpr int RV = S::RV;
#pragma unroll
      for (int ib = 0; ib < N; ++ib)
#pragma unroll
        for (int c = 0; c < N; c += RV)
          {
            const Vec<T, RV> v = *reinterpret_cast<const Vec<T, RV> *>(
              buf + base + ib * PL::W(b) + c);
#pragma unroll
            for (int e = 0; e < RV; ++e)
              u[c + e + N * ib] = v.v[e];
          }
    }
  else
    {
      const volatile __attribute__((address_space(3))) T *vb =
        (const volatile __attribute__((address_space(3))) T *)buf;
#pragma unroll
      for (int r = 0; r < N * N; ++r)
        u[r] = vb[base + (r % N) * PL::W(a) + (r / N) * PL::W(b)];
    }
}

template <int D, int N, typename T, int p>
__device__ inline void tile_write(T *buf, int base, const T *u)
{
  using S = Shape<D, N, T>;
  using PL = Plan<D, N>;
  constexpr int a = PL::a(p), b = PL::b(p);
  if constexpr (a == 0 && S::kRowVec)
    {
      constexpr int RV = S::RV;
#pragma unroll
      for (int ib = 0; ib < N; ++ib)
#pragma unroll
        for (int c = 0; c < N; c += RV)
          {
            Vec<T, RV> v;
#pragma unroll
            for (int e = 0; e < RV; ++e)
              v.v[e] = u[c + e + N * ib];
            *reinterpret_cast<Vec<T, RV> *>(buf + base + ib * PL::W(b) + c) =
              v;
          }
    }
  else
    {
      volatile __attribute__((address_space(3))) T *vb =
        (volatile __attribute__((address_space(3))) T *)buf;
#pragma unroll
      for (int r = 0; r < N * N; ++r)
        vb[base + (r % N) * PL::W(a) + (r / N) * PL::W(b)] = u[r];
    }
}

// Contributions of layout p's directions to its tile: out += A u + b face.
//   u: tile values, r = i_a + N i_b; face[0][l]: inflow value of the line
//   along a with i_b = l; face[1][l]: of the line along b with i_a = l
// read-only data in the constant address space: uniform loads from it are
// scalar loads through the scalar cache
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
// global memory: kernel arguments are known to point there, but a device
// function's pointer arguments are generic (FLAT accesses) unless typed so
template <typename T>
using gptr = __attribute__((address_space(1))) T *;

template <int D, int N, typename T, int p>
__device__ inline void tile_contrib(cptr<T> coef, const T *u,
                                    const T (*face)[N], T *out)
{
  using PL = Plan<D, N>;
  constexpr int NN = N * N;
  constexpr int a = PL::a(p), b = PL::b(p);
  // The coefficients of all directions do not fit in SGPRs: hide the
  // pointer from loop-invariant code motion so that each direction's scalar
  // loads stay next to their use instead of being spilled across the loop.
  asm volatile("" : "+s"(coef));
  {
    cptr<T> A = coef + a * (NN + N);
    cptr<T> bv = A + NN;
#pragma unroll
    for (int l = 0; l < N; ++l)
#pragma unroll
      for (int i = 0; i < N; ++i)
        {
          T acc = bv[i] * face[0][l];
#pragma unroll
          for (int j = 0; j < N; ++j)
            acc += A[i * N + j] * u[j + N * l];
          out[i + N * l] += acc;
        }
  }
  if constexpr (PL::b_active(p))
    {
      asm volatile("" : "+s"(coef));
      cptr<T> A = coef + b * (NN + N);
      cptr<T> bv = A + NN;
#pragma unroll
      for (int l = 0; l < N; ++l)
#pragma unroll
        for (int i = 0; i < N; ++i)
          {
            T acc = bv[i] * face[1][l];
#pragma unroll
            for (int j = 0; j < N; ++j)
              acc += A[i * N + j] * u[l + N * j];
            out[l + N * i] += acc;
          }
    }
}

// the inflow face values of layout p's lines of one cell
//   nb[d]: the inflow neighbour cell across d; fnode[d]: its face node
//   skip_a: the direction-a face will be taken from LDS (a wave-uniform
//   condition): no loads for it
template <int D, int N, typename T, int p>
__device__ inline void face_load(gptr<const T> __restrict__ src, int gbase,
                                 const long long *nb, const int *fnode,
                                 T (*face)[N], bool skip_a = false)
{
  using S = Shape<D, N, T>;
  using PL = Plan<D, N>;
  constexpr int CELL = S::CELL;
  constexpr int a = PL::a(p), b = PL::b(p);
  constexpr int sa = ipow(N, a), sb = ipow(N, b);
  if (!skip_a)
    {
      gptr<const T> s0 = src + nb[a] * CELL + gbase + fnode[a] * sa;
#pragma unroll
      for (int l = 0; l < N; ++l)
        face[0][l] = s0[l * sb];
    }
  if constexpr (PL::b_active(p))
    {
      gptr<const T> s1 = src + nb[b] * CELL + gbase + fnode[b] * sb;
      if constexpr (a == 0 && S::kFaceVec)
        {
          // the N values of the b face along direction 0 are contiguous and
          // start at a multiple of N elements
          constexpr int RV = S::RV;
#pragma unroll
          for (int c = 0; c < N; c += RV)
            {
              gptr<const Vec<T, RV>> v =
                reinterpret_cast<gptr<const Vec<T, RV>>>(s1 + c);
#pragma unroll
              for (int e = 0; e < RV; ++e)
                face[1][c + e] = v->v[e];
            }
        }
      else
        {
#pragma unroll
          for (int l = 0; l < N; ++l)
            face[1][l] = s1[l * sa];
        }
    }
}

// this block's cells (from c0, ncell of them) as 16-byte chunks, chunk
// tid + k BLOCK into ch[k]; indices past the last chunk re-read it
//   groups: the chunks of cell `mycell` (the thread's own), tau + k TPC
template <int D, int N, typename T, int G, int BT, int LB, typename Chunks>
__device__ __forceinline__ void load_chunks(Chunks &ch,
                                            gptr<const T> __restrict__ src,
                                            long long c0, int ncell,
                                            long long mycell, int tau)
{
  using S = Shape<D, N, T, G, BT, LB>;
  constexpr int VEC = S::VEC;
  if constexpr (S::kGroup)
    {
      gptr<const Vec<T, VEC>> g =
        reinterpret_cast<gptr<const Vec<T, VEC>>>(src + mycell * S::CELL);
#pragma unroll
      for (int k = 0; k < S::CPT; ++k)
        {
          gptr<const Vec<T, VEC>> v = g + tau + k * S::TPC;
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            ch[k * VEC + e] = v->v[e];
        }
    }
  else
    {
      gptr<const Vec<T, VEC>> g =
        reinterpret_cast<gptr<const Vec<T, VEC>>>(src + c0 * S::CELL);
      const int last = ncell * S::CELL / VEC - 1;
#pragma unroll
      for (int k = 0; k < S::CPT; ++k)
        {
          gptr<const Vec<T, VEC>> v =
            g + min((int)threadIdx.x + k * S::BLOCK, last);
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            ch[k * VEC + e] = v->v[e];
        }
    }
}

template <int D, int N, typename T, int G, int BT, int LB, typename Chunks>
__device__ __forceinline__ void store_chunks(const Chunks &ch, T *U, int lc,
                                             int tau)
{
  using S = Shape<D, N, T, G, BT, LB>;
  constexpr int VEC = S::VEC, CELL = S::CELL, CELLP = S::CELLP;
#pragma unroll
  for (int k = 0; k < S::CPT; ++k)
    {
      int l, li;
      if constexpr (S::kGroup)
        {
          l = lc;
          li = (tau + k * S::TPC) * VEC;
        }
      else
        {
          const int c = min((int)threadIdx.x + k * S::BLOCK, S::CHUNKS - 1);
          const int i = c * VEC;
          l = i / CELL;
          li = i - l * CELL;
        }
      if constexpr (S::kChunkRuns)
        {
          Vec<T, VEC> v;
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            v.v[e] = ch[k * VEC + e];
          *reinterpret_cast<Vec<T, VEC> *>(U + l * CELLP +
                                           lds_offset<D, N>(li)) = v;
        }
      else
#pragma unroll
        for (int e = 0; e < VEC; ++e)
          U[l * CELLP + lds_offset<D, N>(li + e)] = ch[k * VEC + e];
    }
}

// upwind neighbours of `cell` (cell ids fit in 32 bits; checked on the host)
template <int D>
__device__ inline void inflow_cells(long long cell, const MeshDev &mesh,
                                    const InflowDev &inflow, long long *nb)
{
  unsigned rem = (unsigned)(cell / mesh.inner);
#pragma unroll
  for (int d = 0; d < D; ++d)
    {
      const unsigned n = (unsigned)mesh.n[d];
      const unsigned q = rem / n;
      const unsigned c = rem - q * n;
      rem = q;
      const long long st = mesh.stride[d];
      const long long left = c == 0 ? cell + (long long)(n - 1) * st : cell - st;
      const long long right =
        c == n - 1 ? cell - (long long)(n - 1) * st : cell + st;
      // a direction without transport has zero coefficients: any cell will
      // do, take the left one
      nb[d] = inflow.side[d] < 0 ? right : left;
    }
}

// MODE 0: dst = apply(src).  MODE 1: LSRK stage with K = apply(src),
// sol += b_dt K, ti_out = sol_old + a_dt K.  MODE 2: the stage without
// ti_out (last stage).
// Resident blocks per CU: LDS allows two (FP64) or four (FP32) 3+3 cells
// with their R buffers; the FP32 stage keeps the FP64 solution values in
// registers and runs three to stay clear of spills.  A block of G cells
// counts G times.
template <typename T, int MODE, int G>
constexpr int kBlocksPerCu =
  (sizeof(T) == 8 ? 2 : (MODE == 0 ? 4 : 3)) / G > 0
    ? (sizeof(T) == 8 ? 2 : (MODE == 0 ? 4 : 3)) / G
    : 1;

// Residency of the single-operator kernel.  A full block (kBlock threads, one
// wave per SIMD) runs kBlocksPerCu blocks per CU; a trimmed block runs as
// many as LDS holds, at most that count, and its launch bound asks for the
// waves per SIMD those blocks need (HIP's second launch bound is the minimum
// number of waves per SIMD, which fixes the register budget: 512 / waves).
template <int D, int N, typename T, int MODE, int G>
struct Residency
{
  using S = Shape<D, N, T, G>;
  static constexpr int kFull = kBlocksPerCu<T, MODE, G>;
  static constexpr bool kTrim = !S::kGroup && S::BLOCK < kBlock;
  static constexpr int kLdsBytes =
    S::NBUF * S::CPB * S::CELLP * (int)sizeof(T);
  static constexpr int kLdsBlocks = 160 * 1024 / kLdsBytes;
  static constexpr int kBlocks =
    kTrim ? (kLdsBlocks < kFull ? kLdsBlocks : kFull) : kFull;
  static constexpr int kWavesPerEu =
    kTrim ? (kBlocks * (S::BLOCK / 64) + 3) / 4 : kFull;
  static_assert(kBlocks >= 1 && kWavesPerEu >= 1, "residency");
};

// One block's work on one operator: the logical units lb, lb + slots, ...
// below lend, with the cells in U and the partial sums in R (LDS buffers of
// CPB * CELLP values each).  Logical units: G rows of cells along direction
// 0 (row walk) or groups of CPB cells.  BT: threads per block (see Shape).
template <int D, int N, typename T, int MODE, int G, int BT = 0, int LB = 0>
__device__ __forceinline__ void
apply_block(T *__restrict__ U, T *__restrict__ R, const T *__restrict__ src_,
            T *__restrict__ dst_, const T *__restrict__ coef,
            const MeshDev &mesh, const InflowDev &inflow,
            const StageDev &stage, long long lb, long long lend, int slots)
{
  const gptr<const T> src = (gptr<const T>)src_;
  const gptr<T> dst = (gptr<T>)dst_;
  const gptr<double> sol = (gptr<double>)stage.sol;
  const gptr<T> ti_out = (gptr<T>)static_cast<T *>(stage.ti_out);
  using S = Shape<D, N, T, G, BT, LB>;
  using PL = Plan<D, N>;
  constexpr int P = S::P, NN = S::NN, CELL = S::CELL, TPC = S::TPC;
  constexpr int CPB = S::CPB, CELLP = S::CELLP, VEC = S::VEC, CPT = S::CPT;
  constexpr bool RW = PL::kRowWalk;
  // Large tiles (N >= 5): the last layout reads its tile from U after the
  // other layouts instead of before them, which keeps one tile fewer in
  // registers and costs one more barrier per step.
  constexpr bool kLateLast = P > 1 && N >= 5;
  const cptr<T> ccoef = (cptr<T>)coef;

  const int tid = threadIdx.x;
  // threads past CPB * TPC shadow the last cell: they load and compute but
  // write nothing
  int lc = min(tid / TPC, CPB - 1);
  // TPC is a multiple of the wave size: lc is uniform over a wave, and
  // making that visible keeps the cell arithmetic on the scalar unit
  if constexpr (TPC % 64 == 0)
    lc = __builtin_amdgcn_readfirstlane(lc);
  const int tau = tid % TPC;
  const bool owner = tid < S::THREADS;

  int gbase[P], lbase[P];
  static_for<P>([&](auto pc) {
    constexpr int p = decltype(pc)::value;
    tile_base<D, N>(tau, PL::a(p), PL::b(p), gbase[p], lbase[p]);
    lbase[p] += lc * CELLP;
  });
  int fnode[D];
#pragma unroll
  for (int d = 0; d < D; ++d)
    fnode[d] = inflow.side[d] < 0 ? 0 : N - 1;

  T ch[CPT * VEC];
  T fc[P][2][N];

  // Cells.  Row walk: unit u holds the rows u G .. u G + G - 1 (neighbours
  // along direction 1); at row position k the block works on their cells
  // c0 + l n_0, l < G, with c0 = u G n_0 + x0(k).  Otherwise unit u is the
  // cells u CPB .. u CPB + CPB - 1.
  const int rowlen = RW ? mesh.n[0] : 1;
  const bool fwd = inflow.side[0] >= 0;
  auto first_cell = [&](long long unit, int k) -> long long {
    return RW ? unit * CPB * rowlen + (fwd ? k : rowlen - 1 - k)
              : unit * CPB;
  };
  auto my_cell = [&](long long first) -> long long {
    return RW ? first + (long long)lc * rowlen
              : min(first + lc, mesh.n_cells - 1);
  };
  // direction-1 inflow neighbour inside the group: its face is read from U
  const int nb1 = inflow.side[1] < 0 ? lc + 1 : lc - 1;
  const bool d1_lds = RW && CPB > 1 && nb1 >= 0 && nb1 < CPB;
  constexpr int P1 = RW ? 1 : -1; // the layout with a(p) == 1
  static_assert(!RW || PL::a(1) == 1, "row walk layouts");

  // row walk: the x0 inflow face of the current cell, captured from the
  // previous cell in U (valid when prev_in_row)
  T f0[N];
  bool prev_in_row = false;

  int k = 0;
  long long c0 = first_cell(lb, 0);
  int ncell = (int)min((long long)CPB, mesh.n_cells - c0);
  {
    long long nb[D];
    const long long mc = my_cell(c0);
    inflow_cells<D>(mc, mesh, inflow, nb);
    load_chunks<D, N, T, G, BT, LB>(ch, src, c0, ncell, mc, tau);
    static_for<P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      face_load<D, N, T, p>(src, gbase[p], nb, fnode, fc[p], p == P1 && d1_lds);
    });
  }

  for (;;)
    {
      // 1) this step's cells into U
      store_chunks<D, N, T, G, BT, LB>(ch, U, lc, tau);
      lds_barrier();

      // 2) the next step's values go out now and land during the arithmetic
      // (the last iteration re-reads its own cells); their faces follow
      // layout by layout, as soon as the current layout has consumed its own
      const bool in_row = RW && k + 1 < rowlen;
      const long long ln = in_row ? lb : lb + slots;
      const int nk = in_row ? k + 1 : 0;
      const bool more = ln < lend;
      const long long nc0 = more ? first_cell(ln, nk) : c0;
      const int nncell = (int)min((long long)CPB, mesh.n_cells - nc0);
      const long long nmc = my_cell(nc0);
      load_chunks<D, N, T, G, BT, LB>(ch, src, nc0, nncell, nmc, tau);
      long long nnb[D];
      inflow_cells<D>(nmc, mesh, inflow, nnb);

      // 3) layouts 0 .. P-2 add their contributions into R
      T ul[NN];
      if constexpr (P > 1 && !kLateLast)
        tile_read<D, N, T, P - 1>(U, lbase[P - 1], ul);
      static_for<P - 1>([&](auto pc) {
        constexpr int p = decltype(pc)::value;
        T u[NN], out[NN];
        tile_read<D, N, T, p>(U, lbase[p], u);
        if constexpr (RW && p == 0)
          {
            if (prev_in_row)
#pragma unroll
              for (int l = 0; l < N; ++l)
                fc[p][0][l] = f0[l];
          }
        if constexpr (p == P1)
          {
            // the x1 inflow face of a cell whose x1 neighbour is in the block
            if (d1_lds)
              {
                const volatile __attribute__((address_space(3))) T *vu =
                  (const volatile __attribute__((address_space(3))) T *)U;
                const int base = lbase[p] + (nb1 - lc) * CELLP +
                                 fnode[1] * PL::W(1);
#pragma unroll
                for (int l = 0; l < N; ++l)
                  fc[p][0][l] = vu[base + l * PL::W(PL::b(p))];
              }
          }
        if constexpr (p == 0)
          {
#pragma unroll
            for (int r = 0; r < NN; ++r)
              out[r] = T(0);
          }
        else
          tile_read<D, N, T, p>(R, lbase[p], out);
        tile_contrib<D, N, T, p>(ccoef, u, fc[p], out);
        if (S::kAllOwn || owner)
          tile_write<D, N, T, p>(R, lbase[p], out);
        face_load<D, N, T, p>(src, gbase[p], nnb, fnode, fc[p],
                              (RW && p == 0 && in_row) || (p == P1 && d1_lds));
        if constexpr (RW && p == 0)
          {
            // the next cell's x0 inflow face is this cell's outflow face
            if (in_row)
              {
                const volatile __attribute__((address_space(3))) T *vu =
                  (const volatile __attribute__((address_space(3))) T *)U;
#pragma unroll
                for (int l = 0; l < N; ++l)
                  f0[l] = vu[lbase[p] + fnode[0] * PL::W(0) +
                             l * PL::W(PL::b(p))];
              }
          }
        lds_barrier();
      });

      // 4) the last layout: R + its own contribution, straight to HBM
      // (non-temporal: the output is not read again by this kernel and
      // should not push face lines out of L2)
      {
        constexpr int p = P - 1;
        constexpr int a = PL::a(p), b = PL::b(p);
        const bool store =
          S::kAllOwn && (RW || CPB == 1) ? true : (owner && lc < ncell);
        const long long g0 = my_cell(c0) * CELL + gbase[p];
        // the stage's solution values go out before the arithmetic (small
        // tiles; large ones load them at the store)
        constexpr bool kEarlySol = MODE != 0 && NN <= 16;
        double sv[MODE == 0 ? 1 : NN];
        if constexpr (kEarlySol)
          {
#pragma unroll
            for (int r = 0; r < NN; ++r)
              sv[r] = sol[g0 + (r % N) * ipow(N, a) + (r / N) * ipow(N, b)];
          }
        T out[NN];
        if constexpr (kLateLast)
          tile_read<D, N, T, p>(U, lbase[p], ul);
        if constexpr (P > 1)
          tile_read<D, N, T, p>(R, lbase[p], out);
        else
          {
            tile_read<D, N, T, p>(U, lbase[p], ul);
#pragma unroll
            for (int r = 0; r < NN; ++r)
              out[r] = T(0);
          }
        tile_contrib<D, N, T, p>(ccoef, ul, fc[p], out);
        face_load<D, N, T, p>(src, gbase[p], nnb, fnode, fc[p]);
        if (store)
          {
            if constexpr (MODE != 0 && !kEarlySol)
              {
#pragma unroll
                for (int r = 0; r < NN; ++r)
                  sv[r] =
                    sol[g0 + (r % N) * ipow(N, a) + (r / N) * ipow(N, b)];
              }
#pragma unroll
            for (int r = 0; r < NN; ++r)
              {
                const long long g =
                  g0 + (r % N) * ipow(N, a) + (r / N) * ipow(N, b);
                if constexpr (MODE == 0)
                  __builtin_nontemporal_store(out[r], &dst[g]);
                else
                  {
                    __builtin_nontemporal_store(
                      sv[r] + stage.b_dt * (double)out[r], &sol[g]);
                    if constexpr (MODE == 1)
                      __builtin_nontemporal_store(
                        (T)(sv[r] + stage.a_dt * (double)out[r]),
                        &ti_out[g]);
                  }
              }
          }
      }
      if (!more)
        break;
      lb = ln;
      k = nk;
      c0 = nc0;
      ncell = nncell;
      prev_in_row = in_row;
      // one layout, or the last layout's late tile read: this iteration's
      // U reads precede the next step's chunks
      if constexpr (P == 1 || kLateLast)
        lds_barrier();
    }
}

template <int D, int N, typename T, int MODE, int G>
__global__ __launch_bounds__((Shape<D, N, T, G>::BLOCK),
                             (Residency<D, N, T, MODE, G>::kWavesPerEu)) void
  apply_kernel(
  const T *__restrict__ src, T *__restrict__ dst, const T *__restrict__ coef,
  MeshDev mesh, InflowDev inflow, StageDev stage, long long n_logical,
  int slots)
{
  using S = Shape<D, N, T, G>;
  __shared__ __align__(16) T U[S::CPB * S::CELLP];
  __shared__ __align__(16) T R[S::P > 1 ? S::CPB * S::CELLP : 1];
  // this block's share of its XCD's contiguous range of logical units
  const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
  const long long per = (n_logical + 7) / 8;
  const long long lend = min(n_logical, (xcd + 1) * per);
  const long long lb = xcd * per + slot;
  if (lb >= lend)
    return; // uniform for the block, before any barrier
  apply_block<D, N, T, MODE, G>(U, R, src, dst, coef, mesh, inflow, stage, lb,
                                lend, slots);
}

int g_cus = 0;

int cus_per_xcd()
{
  if (g_cus == 0)
    {
      int dev = 0, cus = 256;
      if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus,
                                    hipDeviceAttributeMultiprocessorCount, dev);
      g_cus = (cus + 7) / 8;
    }
  return g_cus;
}

template <int D, int N, typename T, int G>
hipError_t launch_g(int mode, const T *src, T *dst, const T *coef,
                    const MeshDev &mesh, const InflowDev &inflow,
                    const StageDev &stage, hipStream_t stream)
{
  using S = Shape<D, N, T, G>;
  constexpr bool RW = Plan<D, N>::kRowWalk;
  const long long n_logical =
    RW ? mesh.n_cells / ((long long)mesh.n[0] * G)
       : (mesh.n_cells + S::CPB - 1) / S::CPB;
  const long long per = (n_logical + 7) / 8;
  auto go = [&](auto kernel, int bpc) {
    // Row walk: one block per unit, so the dispatcher hands the rows of an
    // XCD out in order and the blocks resident on it work on neighbouring
    // rows (their x2 / x3 inflow faces are lines another block has just
    // loaded into the XCD's L2).  Otherwise persistent blocks, no more per
    // XCD than it has units.
    const int used =
      RW ? (int)per : (int)min((long long)(bpc * cus_per_xcd()), per);
    hipLaunchKernelGGL(kernel, dim3(8u * (unsigned)used), dim3(S::BLOCK), 0,
                       stream, src, dst, coef, mesh, inflow, stage, n_logical,
                       used);
  };
  const int m = mode == 0 ? 0 : (stage.ti_out ? 1 : 2);
  if (m == 0)
    go(apply_kernel<D, N, T, 0, G>, Residency<D, N, T, 0, G>::kBlocks);
  else if (m == 1)
    go(apply_kernel<D, N, T, 1, G>, Residency<D, N, T, 1, G>::kBlocks);
  else
    go(apply_kernel<D, N, T, 2, G>, Residency<D, N, T, 2, G>::kBlocks);
  return hipGetLastError();
}

// Cells per block along direction 1 for the row walk.  Groups of two (FP64)
// or four (FP32) read the x1 inflow face of all but one cell from LDS, but on
// MI355X at 3+3, k=3, 8^6 cells they measured no faster (FP64 251 vs 252
// GDoF/s) or slower (FP32, four cells of 1024 threads: 419 vs 485), so one
// cell per block is the default; the group size stays a tuning parameter.
template <typename T>
int preferred_group(int mode)
{
  (void)mode;
  return 1;
}

template <int D, int N, typename T>
hipError_t launch(int mode, const T *src, T *dst, const T *coef,
                  const MeshDev &mesh, const InflowDev &inflow,
                  const StageDev &stage, hipStream_t stream)
{
  if constexpr (Plan<D, N>::kRowWalk)
    {
      int g = preferred_group<T>(mode);
      while (g > 1 && mesh.n[1] % g != 0)
        g /= 2;
      if constexpr (sizeof(T) == 4)
        if (g == 4)
          return launch_g<D, N, T, 4>(mode, src, dst, coef, mesh, inflow,
                                      stage, stream);
      if (g == 2)
        return launch_g<D, N, T, 2>(mode, src, dst, coef, mesh, inflow, stage,
                                    stream);
    }
  return launch_g<D, N, T, 1>(mode, src, dst, coef, mesh, inflow, stage,
                              stream);
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

// Batches: one kernel per degree N in the batch, with the block shape and
// residency of the single-operator kernel of that degree (a thread owns an
// N x N tile, the block holds as many cells as LDS and the register budget
// allow).  A block walks its tasks -- ranges of logical units of one
// operator of degree N -- in order, each through the same block code as the
// single-operator kernel, so a batched apply or stage equals the
// single-operator call bit for bit.  The operator's mesh and inflow sides
// are read field by field from the device table (uniform scalar loads).
template <int N, typename T, int MODE>
__global__ __launch_bounds__((Shape<4, N, T, 1>::BLOCK),
                             (Residency<4, N, T, MODE, 1>::kWavesPerEu)) void
  batch_kernel(const BatchOp *__restrict__ ops,
               const BatchTask *__restrict__ tasks, BatchPlan plan,
               BatchPtrs ptrs)
{
  using S = Shape<4, N, T, 1>;
  __shared__ __align__(16) T U[S::CPB * S::CELLP];
  __shared__ __align__(16) T R[S::P > 1 ? S::CPB * S::CELLP : 1];
  // this block's tasks: its strided share of each operator fragment its
  // XCD walks
  const int b = (blockIdx.x % 8) * plan.per + blockIdx.x / 8;
  const int t_end = plan.block_start[b + 1];
  for (int t = plan.block_start[b]; t < t_end; ++t)
    {
      const gptr<const BatchTask> tk = (gptr<const BatchTask>)(tasks + t);
      const int op = tk->op;
      const gptr<const BatchOp> o = (gptr<const BatchOp>)(ops + op);
      MeshDev mesh;
      InflowDev inflow;
      mesh.dim = o->mesh.dim;
      mesh.n_cells = o->mesh.n_cells;
      mesh.inner = o->mesh.inner;
#pragma unroll
      for (int d = 0; d < kMaxDim; ++d)
        {
          mesh.n[d] = o->mesh.n[d];
          mesh.stride[d] = o->mesh.stride[d];
          mesh.lo[d] = o->mesh.lo[d];
          mesh.h[d] = o->mesh.h[d];
          inflow.side[d] = o->inflow.side[d];
        }
      StageDev stage{};
      stage.a_dt = ptrs.a_dt;
      stage.b_dt = ptrs.b_dt;
      stage.sol = MODE == 0 ? nullptr : ptrs.sol[op];
      stage.ti_out = MODE == 1 ? ptrs.ti[op] : nullptr;
      apply_block<4, N, T, MODE, 1>(
        U, R, static_cast<const T *>(ptrs.src[op]),
        MODE == 0 ? static_cast<T *>(ptrs.dst[op]) : nullptr,
        static_cast<const T *>(o->coef), mesh, inflow, stage, tk->lb, tk->le,
        tk->stride);
      // the next task starts writing U: every read of this task's buffers
      // comes first
      lds_barrier();
    }
}

template <int N, typename T>
hipError_t batch_launch(int mode, const BatchOp *ops, const BatchTask *tasks,
                        const BatchPlan &plan, const BatchPtrs &ptrs,
                        hipStream_t stream)
{
  if (plan.per < 1 || !plan.block_start)
    return hipErrorInvalidValue;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(8 * plan.per),
                       dim3(Shape<4, N, T, 1>::BLOCK), 0, stream, ops, tasks,
                       plan, ptrs);
  };
  if (mode == 0)
    go(batch_kernel<N, T, 0>);
  else if (mode == 1)
    go(batch_kernel<N, T, 1>);
  else
    go(batch_kernel<N, T, 2>);
  return hipGetLastError();
}

// dst[op] = (T) sol[op] for every operator: grid row op converts operator
// op, four values per thread and step (two 16-byte loads, 16 or 32 bytes
// stored), the last n % 4 values by block 0
template <typename T>
__global__ void batch_convert_kernel(BatchPtrs ptrs, BatchSizes sizes)
{
  const int op = blockIdx.y;
  const long long n = sizes.off[op + 1] - sizes.off[op];
  const double *__restrict__ src = ptrs.sol[op];
  T *__restrict__ dst = static_cast<T *>(ptrs.dst[op]);
  const long long n4 = n / 4;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n4;
       q += (long long)gridDim.x * blockDim.x)
    {
      const Vec<double, 2> a = reinterpret_cast<const Vec<double, 2> *>(src)[2 * q];
      const Vec<double, 2> b =
        reinterpret_cast<const Vec<double, 2> *>(src)[2 * q + 1];
      Vec<T, 4> o;
      o.v[0] = (T)a.v[0];
      o.v[1] = (T)a.v[1];
      o.v[2] = (T)b.v[0];
      o.v[3] = (T)b.v[1];
      reinterpret_cast<Vec<T, 4> *>(dst)[q] = o;
    }
  if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4)
    dst[4 * n4 + threadIdx.x] = (T)src[4 * n4 + threadIdx.x];
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
bool collapsed_batch_supported(int D, int N)
{
  return D == 4 && N >= 2 && N <= 6;
}

namespace
{
template <typename T>
int batch_cpb(int N)
{
  using collapsed::Shape;
  switch (N)
    {
      case 2:
        return Shape<4, 2, T, 1>::CPB;
      case 3:
        return Shape<4, 3, T, 1>::CPB;
      case 4:
        return Shape<4, 4, T, 1>::CPB;
      case 5:
        return Shape<4, 5, T, 1>::CPB;
      default:
        return Shape<4, 6, T, 1>::CPB;
    }
}

// resident blocks per CU of the stage kernel of degree N
template <typename T>
int batch_bpc(int N)
{
  using collapsed::Residency;
  switch (N)
    {
      case 2:
        return Residency<4, 2, T, 1, 1>::kBlocks;
      case 3:
        return Residency<4, 3, T, 1, 1>::kBlocks;
      case 4:
        return Residency<4, 4, T, 1, 1>::kBlocks;
      case 5:
        return Residency<4, 5, T, 1, 1>::kBlocks;
      default:
        return Residency<4, 6, T, 1, 1>::kBlocks;
    }
}
} // namespace

// degrees with the wave-per-group kernel (hd_cvwave.hip) batch through it,
// so that a batch equals its single-operator calls bit for bit
int collapsed_batch_cells_per_block(int D, int N, int precision)
{
  if (!collapsed_batch_supported(D, N))
    return 0;
  if (collapsed_wave_supported(D, N, precision))
    return collapsed_wave_cells_per_unit(N);
  return precision == 0 ? batch_cpb<double>(N) : batch_cpb<float>(N);
}

int collapsed_batch_blocks_per_xcd(int N, int precision)
{
  if (N < 2 || N > 6)
    return 0;
  if (collapsed_wave_supported(4, N, precision))
    return collapsed_wave_blocks_per_cu(N, precision) *
           collapsed::cus_per_xcd();
  return (precision == 0 ? batch_bpc<double>(N) : batch_bpc<float>(N)) *
         collapsed::cus_per_xcd();
}

hipError_t collapsed_batch(int D, int N, int precision, int mode,
                           const BatchOp *ops, const BatchTask *tasks,
                           const BatchPlan &plan, const BatchPtrs &ptrs,
                           hipStream_t stream)
{
  if (!collapsed_batch_supported(D, N))
    return hipErrorInvalidValue;
  if (collapsed_wave_supported(D, N, precision))
    return collapsed_wave_batch(N, precision, mode, ops, tasks, plan, ptrs,
                                stream);
#define HD_BN(NN_)                                                           \
  if (N == NN_)                                                              \
    return precision == 0                                                    \
             ? collapsed::batch_launch<NN_, double>(mode, ops, tasks, plan,  \
                                                    ptrs, stream)            \
             : collapsed::batch_launch<NN_, float>(mode, ops, tasks, plan,   \
                                                   ptrs, stream);
  HD_BN(2) HD_BN(3) HD_BN(4) HD_BN(5) HD_BN(6)
#undef HD_BN
  return hipErrorInvalidValue;
}

hipError_t batch_convert(int precision, int n_ops, const long long *n_dofs,
                         const BatchPtrs &ptrs, hipStream_t stream)
{
  BatchSizes sizes{};
  sizes.n = n_ops;
  long long most = 0;
  for (int i = 0; i < n_ops; ++i)
    {
      sizes.off[i + 1] = sizes.off[i] + n_dofs[i];
      most = std::max(most, n_dofs[i]);
      // 16-byte accesses: every vector starts 16-byte aligned
      if (reinterpret_cast<uintptr_t>(ptrs.sol[i]) % 16 != 0 ||
          reinterpret_cast<uintptr_t>(ptrs.dst[i]) % 16 != 0)
        return hipErrorInvalidValue;
    }
  if (n_ops == 0 || most == 0)
    return hipSuccess;
  const dim3 grid((unsigned)std::min<long long>((most / 4 + 255) / 256, 1024),
                  (unsigned)n_ops);
  if (precision == 0)
    hipLaunchKernelGGL(collapsed::batch_convert_kernel<double>, grid,
                       dim3(256), 0, stream, ptrs, sizes);
  else
    hipLaunchKernelGGL(collapsed::batch_convert_kernel<float>, grid,
                       dim3(256), 0, stream, ptrs, sizes);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Split form for cells larger than LDS (3+3 at k = 5: 6^6 values, 373 KB in
// FP64).  The collapsed operator is a sum of 1D operators, one per
// direction, so it splits exactly:
//   part = sum_{d < 4} (A_d u along d + b_d face_d)
//          the 4D kernel, unchanged, on the slabs of fixed (i4, ..): a slab
//          is N^4 contiguous values and its neighbour across d < 4 is the
//          same slab of the neighbouring cell, so the slabs form a 4D mesh
//          with strides scaled by the slab count (MeshDev::inner);
//   K    = part + sum_{d >= 4} (A_d u along d + b_d face_d)
//          the plane kernel: one thread per (cell, point of the first four
//          directions) holds the N^{D-4} values of its plane in registers,
//          adds the part and finishes the apply or the LSRK stage.
// Traffic: the slab pass reads u and writes part, the plane pass reads u,
// part and the faces across d >= 4 and writes the result (FP64 apply about
// 48 bytes per DoF instead of 16).
namespace collapsed
{
// three waves per SIMD: with the part loaded up front the compiler would
// otherwise keep every plane value live and run one wave per SIMD
template <int D, int N, typename T, int MODE>
__global__ __launch_bounds__(256, 3) void plane_kernel(
  const T *__restrict__ src, const T *part, T *dst,
  const T *__restrict__ coef, MeshDev mesh, InflowDev inflow, StageDev stage)
{
  constexpr int R = D - 4, P = ipow(N, 4), PL = ipow(N, R);
  constexpr long long CELL = (long long)P * PL;
  constexpr int NN = N * N;
  const cptr<T> cc = (cptr<T>)coef;
  const long long n = mesh.n_cells * P;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       idx < n; idx += (long long)gridDim.x * blockDim.x)
    {
      const long long cell = idx / P;
      const int pt = (int)(idx - cell * P);
      const long long g0 = cell * CELL + pt;
      // the part of the first four directions starts the sums: its loads
      // go out with the plane's values (one round trip per plane)
      T u[PL], k[PL];
#pragma unroll
      for (int q = 0; q < PL; ++q)
        {
          u[q] = src[g0 + (long long)q * P];
          k[q] = part[g0 + (long long)q * P];
        }
#pragma unroll
      for (int r = 0; r < R; ++r)
        {
          const int d = 4 + r;
          const int sd = r == 0 ? 1 : N; // plane stride of direction d
          const int side = inflow.side[d];
          const long long nb = inflow_neighbour(mesh, cell, d, side);
          const int fnode = side < 0 ? 0 : N - 1;
          const T *fsrc = src + nb * CELL + pt;
          const cptr<T> A = cc + d * (NN + N);
          const cptr<T> bv = A + NN;
#pragma unroll
          for (int l = 0; l < PL / N; ++l)
            {
              // base of line l along d in the plane
              const int base = r == 0 ? l * N : l;
              const T fv = fsrc[(long long)(base + fnode * sd) * P];
#pragma unroll
              for (int i = 0; i < N; ++i)
                {
                  T acc = bv[i] * fv;
#pragma unroll
                  for (int j = 0; j < N; ++j)
                    acc += A[i * N + j] * u[base + j * sd];
                  k[base + i * sd] += acc;
                }
            }
        }
#pragma unroll
      for (int q = 0; q < PL; ++q)
        {
          const long long g = g0 + (long long)q * P;
          const T kq = k[q];
          if constexpr (MODE == 0)
            dst[g] = kq;
          else
            {
              const double so = stage.sol[g];
              stage.sol[g] = so + stage.b_dt * (double)kq;
              if constexpr (MODE == 1)
                static_cast<T *>(stage.ti_out)[g] =
                  (T)(so + stage.a_dt * (double)kq);
            }
        }
    }
}

template <int D, int N, typename T>
hipError_t split_launch(int mode, const T *src, T *dst, T *part,
                        const T *coef, const MeshDev &mesh,
                        const InflowDev &inflow, const StageDev &stage,
                        hipStream_t stream)
{
  static_assert(D == 5 || D == 6, "split form: D = 5, 6");
  constexpr int R = D - 4;
  // 1) the slab view: a 4D mesh of n_cells * N^R slabs
  MeshDev m4 = mesh;
  m4.dim = 4;
  m4.inner = ipow(N, R);
  for (int d = 0; d < 4; ++d)
    m4.stride[d] = mesh.stride[d] * m4.inner;
  for (int d = 4; d < kMaxDim; ++d)
    {
      m4.n[d] = 1;
      m4.stride[d] = 0;
    }
  m4.n_cells = mesh.n_cells * m4.inner;
  InflowDev in4 = inflow;
  for (int d = 4; d < kMaxDim; ++d)
    in4.side[d] = 0;
  const StageDev none{};
  hipError_t e =
    launch<4, N, T>(0, src, part, coef, m4, in4, none, stream);
  if (e != hipSuccess)
    return e;
  // 2) the planes
  const long long n = mesh.n_cells * ipow(N, 4);
  const unsigned blocks =
    (unsigned)std::min<long long>((n + 255) / 256, 256LL * 16);
  if (mode == 0)
    hipLaunchKernelGGL((plane_kernel<D, N, T, 0>), dim3(blocks), dim3(256), 0,
                       stream, src, part, dst, coef, mesh, inflow, stage);
  else if (stage.ti_out)
    hipLaunchKernelGGL((plane_kernel<D, N, T, 1>), dim3(blocks), dim3(256), 0,
                       stream, src, part, dst, coef, mesh, inflow, stage);
  else
    hipLaunchKernelGGL((plane_kernel<D, N, T, 2>), dim3(blocks), dim3(256), 0,
                       stream, src, part, dst, coef, mesh, inflow, stage);
  return hipGetLastError();
}

template <typename T>
hipError_t split_dispatch(int D, int N, int mode, const T *src, T *dst,
                          T *part, const T *coef, const MeshDev &mesh,
                          const InflowDev &inflow, const StageDev &stage,
                          hipStream_t stream)
{
  if (D == 6 && N == 5)
    return split_launch<6, 5, T>(mode, src, dst, part, coef, mesh, inflow,
                                 stage, stream);
  if (D == 6 && N == 6)
    return split_launch<6, 6, T>(mode, src, dst, part, coef, mesh, inflow,
                                 stage, stream);
  if (D == 5 && N == 5)
    return split_launch<5, 5, T>(mode, src, dst, part, coef, mesh, inflow,
                                 stage, stream);
  if (D == 5 && N == 6)
    return split_launch<5, 6, T>(mode, src, dst, part, coef, mesh, inflow,
                                 stage, stream);
  return hipErrorInvalidValue;
}
} // namespace collapsed

bool collapsed_split_supported(int D, int N)
{
  return (D == 5 || D == 6) && (N == 5 || N == 6);
}

hipError_t collapsed_split_f64(int D, int N, int mode, const double *src,
                               double *dst, double *part, const double *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream)
{
  return collapsed::split_dispatch<double>(D, N, mode, src, dst, part, coef,
                                           mesh, inflow, stage, stream);
}

hipError_t collapsed_split_f32(int D, int N, int mode, const float *src,
                               float *dst, float *part, const float *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream)
{
  return collapsed::split_dispatch<float>(D, N, mode, src, dst, part, coef,
                                          mesh, inflow, stage, stream);
}
} // namespace hd
