// hd_cres.hip -- the FP32 collapsed apply and LSRK stage of 3+3 cells that
// do not fit LDS (k = 4, 5: 5^6 and 6^6 values), in one pass (gfx950).
//
// The collapsed operator (section "Collapsed apply" of hd_collapsed.hip) is
// a sum of 1D operators, one per direction.  One block per cell: thread t
// owns the 4D points p = t and t + TH of the first four directions and
// keeps the outputs of all (i4, i5) at its points in registers (2 N^2
// accumulators: the cell's result lives in the block's registers, 187 KB at
// k=5 in FP32).  The input streams through LDS one chunk of fixed i5 at a
// time (N^5 values, double-buffered); per chunk j a thread adds
//   - along i5: A5[i5][j] u(p, i4, j) to every output plane i5,
//   - along i4: A4 u(p, ., j) to plane j,
//   - along d < 4: the 1D operator of direction d on the LDS lines through p
//     (per-lane rows of the coefficient table, kept in LDS), to plane j,
// and the inflow face terms b_d f_d (the neighbour's face values from HBM /
// L2).  Every value of the cell is read from HBM once and written once: 8
// bytes per DoF for the FP32 apply plus the faces, against about 24 for the
// two-pass split form.  FP64 does not fit (373 KB of accumulators per
// cell), so FP64 keeps the split form.
#include "hd_device.h"

#include <algorithm>

namespace hd
{
namespace cres
{
template <class T>
using cptr = const __attribute__((address_space(4))) T *;

template <int N>
struct Geo
{
  static constexpr int NN = N * N;
  static constexpr int P = NN * NN;              // 4D points per cell
  static constexpr int TH = (P + 1) / 2;         // threads owning points
  static constexpr int BLOCK = (TH + 63) / 64 * 64;
  static constexpr int CH = P * N;               // values per chunk (i5)
  static constexpr int FC = P;                   // face values per d < 4
  static constexpr int BUF = CH + 4 * FC;        // LDS floats per chunk
  static constexpr long long CELL = (long long)CH * N;
  static constexpr int LPT = (CH + BLOCK - 1) / BLOCK;     // chunk loads
};

// 4 bytes per lane from the lane's own global address into LDS at
// dst_wave + lane (dst_wave uniform over the wave): the data goes straight
// to LDS without registers and lands by the next barrier
__device__ inline void load_lds(const float *g, float *dst_wave)
{
  __builtin_amdgcn_global_load_lds(
    (const __attribute__((address_space(1))) void *)g,
    (__attribute__((address_space(3))) void *)dst_wave, 4, 0, 0);
}

template <int N, int MODE>
__global__ __launch_bounds__(Geo<N>::BLOCK) void cres_kernel(
  const float *__restrict__ src, float *__restrict__ dst,
  const float *__restrict__ coef, MeshDev mesh, InflowDev inflow,
  StageDev stage)
{
  using G = Geo<N>;
  constexpr int NN = G::NN, P = G::P, TH = G::TH, CH = G::CH;
  constexpr int LPT = G::LPT, BLOCK = G::BLOCK;
  constexpr int FC = G::FC, BUF = G::BUF, PN = P / N;
  constexpr long long CELL = G::CELL;
  // per chunk: the N^5 values, then the inflow faces of directions 0..3
  // (face d: the neighbour's values at its face node, indexed by the point
  // without digit d and by i4)
  __shared__ __align__(16) float buf[2][BUF];
  __shared__ float Ad[4][NN + N]; // A_d (rows: output node) and b_d, d < 4
  const cptr<float> cc = (cptr<float>)coef;
  for (int e = threadIdx.x; e < 4 * (NN + N); e += BLOCK)
    (&Ad[0][0])[e] = coef[e];

  constexpr int sd[4] = {1, N, NN, NN * N};
  // the thread index and what derives from it are recomputed from an
  // opaque copy per chunk: no lane-dependent addresses held in registers
  // across the chunk loop
  int t;
  bool own[2];
  int pp[2];
  auto lane_setup = [&]() {
    t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
#pragma unroll
    for (int q = 0; q < 2; ++q)
      {
        const int p = t + q * TH;
        own[q] = t < TH && p < P;
        pp[q] = own[q] ? p : 0;
      }
  };
  lane_setup();
  // inflow face nodes (the neighbour's last node from the left, first from
  // the right)
  int fnode[6];
#pragma unroll
  for (int d = 0; d < 6; ++d)
    fnode[d] = inflow.side[d] < 0 ? 0 : N - 1;
  // coefficients of the last two directions (uniform: scalar loads)
  auto A4 = [&](int i, int k) -> float { return cc[4 * (NN + N) + i * N + k]; };
  auto B4 = [&](int i) -> float { return cc[4 * (NN + N) + NN + i]; };
  auto A5 = [&](int i, int k) -> float { return cc[5 * (NN + N) + i * N + k]; };
  auto B5 = [&](int i) -> float { return cc[5 * (NN + N) + NN + i]; };
  __syncthreads();

  // the inflow neighbours of a cell (uniform over the block)
  struct Nb
  {
    long long d0, d1, d2, d3, d4, d5;
  };
  auto neighbours = [&](long long cl) -> Nb {
    Nb r;
    r.d0 = inflow_neighbour(mesh, cl, 0, inflow.side[0]);
    r.d1 = inflow_neighbour(mesh, cl, 1, inflow.side[1]);
    r.d2 = inflow_neighbour(mesh, cl, 2, inflow.side[2]);
    r.d3 = inflow_neighbour(mesh, cl, 3, inflow.side[3]);
    r.d4 = inflow_neighbour(mesh, cl, 4, inflow.side[4]);
    r.d5 = inflow_neighbour(mesh, cl, 5, inflow.side[5]);
    return r;
  };
  // chunk j of a cell and its faces into LDS buffer b, straight from HBM /
  // L2 (element e of a wave lands at the wave's base + lane).  Face d of
  // the chunk: the neighbour's values at its face node, indexed by the point
  // without digit d and by i4; one direction per pass, so that no
  // lane-dependent choice of neighbour is needed
  constexpr int FPD = (FC + BLOCK - 1) / BLOCK;
  auto fetch = [&](long long cl, const Nb &nbx, int j, int b) {
    const int wbase = t & ~63;
    const float *uc = src + cl * CELL;
#pragma unroll
    for (int l = 0; l < LPT; ++l)
      {
        const int e = t + BLOCK * l;
        if (e < CH)
          load_lds(uc + (long long)CH * j + e, &buf[b][wbase + BLOCK * l]);
      }
#pragma unroll
    for (int d = 0; d < 4; ++d)
      {
        const long long nbd = d == 0   ? nbx.d0
                              : d == 1 ? nbx.d1
                              : d == 2 ? nbx.d2
                                       : nbx.d3;
        const float *fsrc =
          src + nbd * CELL + (long long)fnode[d] * sd[d] + (long long)CH * j;
#pragma unroll
        for (int l = 0; l < FPD; ++l)
          {
            const int e = t + BLOCK * l;
            if (e < FC)
              {
                const int c = e % PN, i4 = e / PN;
                const int pf = c % sd[d] + sd[d] * N * (c / sd[d]);
                load_lds(fsrc + pf + (long long)P * i4,
                         &buf[b][CH + d * FC + wbase + BLOCK * l]);
              }
          }
      }
  };

  // the first cell's first chunk; every later cell's first chunk goes out
  // before the previous cell's outputs
  Nb nb{};
  if ((long long)blockIdx.x < mesh.n_cells)
    {
      nb = neighbours(blockIdx.x);
      fetch(blockIdx.x, nb, 0, 0);
    }
  for (long long cell = blockIdx.x; cell < mesh.n_cells; cell += gridDim.x)
    {
      float acc[2][N][N]; // [point][i4][i5]
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int a = 0; a < N; ++a)
#pragma unroll
          for (int b = 0; b < N; ++b)
            acc[q][a][b] = 0.0f;
      // the i4 and i5 inflow faces of this thread's points go out now and
      // land with the first chunk (no global load inside the chunk loop)
      float f4[2][N], f5[2][N];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int k = 0; k < N; ++k)
          {
            f4[q][k] = src[nb.d4 * CELL + pp[q] + (long long)P * (fnode[4] + N * k)];
            f5[q][k] = src[nb.d5 * CELL + pp[q] + (long long)P * (k + N * fnode[5])];
          }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < N; ++j)
        {
          // the next chunk travels into the other buffer meanwhile (every
          // thread finished reading it before the barrier that ended the
          // previous chunk)
          lane_setup();
          if (j + 1 < N)
            fetch(cell, nb, j + 1, (j + 1) & 1);
          const float *ch = buf[j & 1];
#pragma unroll
          for (int q = 0; q < 2; ++q)
            {
              if (!own[q])
                continue;
              const int p = pp[q];
              float v[N];
#pragma unroll
              for (int k = 0; k < N; ++k)
                v[k] = ch[p + P * k];
              // along i5: plane j of the input feeds every output plane
#pragma unroll
              for (int i5 = 0; i5 < N; ++i5)
                {
                  const float a5 = A5(i5, j);
#pragma unroll
                  for (int i4 = 0; i4 < N; ++i4)
                    acc[q][i4][i5] += a5 * v[i4];
                }
              // along i4, within plane j, with the i4 inflow face
#pragma unroll
              for (int i4 = 0; i4 < N; ++i4)
                {
                  float a = B4(i4) * f4[q][j];
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    a += A4(i4, k) * v[k];
                  acc[q][i4][j] += a;
                }
              // along d < 4: the lines through p in the LDS chunk
#pragma unroll
              for (int d = 0; d < 4; ++d)
                {
                  const int row = (p / sd[d]) % N;
                  const int base = p - row * sd[d];
                  const float *ar = &Ad[d][row * N];
                  const float br = Ad[d][NN + row];
                  // face index of p in direction d: p without its digit d
                  const int pc = p % sd[d] + sd[d] * (p / (sd[d] * N));
                  const float *fd = ch + CH + d * FC + pc;
#pragma unroll
                  for (int i4 = 0; i4 < N; ++i4)
                    {
                      float a = br * fd[PN * i4];
#pragma unroll
                      for (int k = 0; k < N; ++k)
                        a += ar[k] * ch[base + k * sd[d] + P * i4];
                      acc[q][i4][j] += a;
                    }
                }
            }
          // the next chunk's loads have landed and this chunk's reads are
          // done (the barrier waits for both)
          __syncthreads();
        }
      // the next cell's first chunk travels while this cell's outputs go
      // out (both buffers are free after the last chunk's barrier); then
      // the outputs with the i5 inflow face (for fixed (i4, i5) the points
      // of consecutive threads are consecutive values)
      const long long next = cell + gridDim.x;
      if (next < mesh.n_cells)
        {
          lane_setup();
          nb = neighbours(next);
          fetch(next, nb, 0, 0);
        }
#pragma unroll
      for (int q = 0; q < 2; ++q)
        {
          if (!own[q])
            continue;
          const int p = pp[q];
          const long long g0 = cell * CELL + p;
#pragma unroll
          for (int i4 = 0; i4 < N; ++i4)
            {
              // a stage's solution values of this i4 row: N loads in flight
              // before the first store
              double so[MODE == 0 ? 1 : N];
              if constexpr (MODE != 0)
#pragma unroll
                for (int i5 = 0; i5 < N; ++i5)
                  so[i5] = stage.sol[g0 + (long long)P * (i4 + N * i5)];
#pragma unroll
              for (int i5 = 0; i5 < N; ++i5)
                {
                  const float k = acc[q][i4][i5] + B5(i5) * f5[q][i4];
                  const long long g = g0 + (long long)P * (i4 + N * i5);
                  if constexpr (MODE == 0)
                    dst[g] = k;
                  else
                    {
                      stage.sol[g] = so[i5] + stage.b_dt * (double)k;
                      if constexpr (MODE == 1)
                        static_cast<float *>(stage.ti_out)[g] =
                          (float)(so[i5] + stage.a_dt * (double)k);
                    }
                }
            }
        }
    }
}

template <int N>
hipError_t launch(int mode, const float *src, float *dst, const float *coef,
                  const MeshDev &mesh, const InflowDev &inflow,
                  const StageDev &stage, hipStream_t stream)
{
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  const unsigned grid =
    (unsigned)std::min<long long>(mesh.n_cells, (long long)cus);
  const dim3 block(Geo<N>::BLOCK);
  if (mode == 0)
    hipLaunchKernelGGL((cres_kernel<N, 0>), dim3(grid), block, 0, stream,
                       src, dst, coef, mesh, inflow, stage);
  else if (stage.ti_out)
    hipLaunchKernelGGL((cres_kernel<N, 1>), dim3(grid), block, 0, stream,
                       src, dst, coef, mesh, inflow, stage);
  else
    hipLaunchKernelGGL((cres_kernel<N, 2>), dim3(grid), block, 0, stream,
                       src, dst, coef, mesh, inflow, stage);
  return hipGetLastError();
}
} // namespace cres

bool collapsed_cres_supported(int D, int N)
{
  return D == 6 && (N == 5 || N == 6);
}

hipError_t collapsed_cres_f32(int N, int mode, const float *src, float *dst,
                              const float *coef, const MeshDev &mesh,
                              const InflowDev &inflow, const StageDev &stage,
                              hipStream_t stream)
{
  if (N == 5)
    return cres::launch<5>(mode, src, dst, coef, mesh, inflow, stage, stream);
  if (N == 6)
    return cres::launch<6>(mode, src, dst, coef, mesh, inflow, stage, stream);
  return hipErrorInvalidValue;
}
} // namespace hd
