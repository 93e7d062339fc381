// hd_cvwave.hip -- the 2+2 constant-velocity collapsed apply and fused LSRK
// stage, one wave per group of cells (gfx950).
//
// The operator (the collapsed form of advection_operation.h:221-566 for a
// constant velocity on an affine Cartesian mesh):
//   K(p) = sum_d [ sum_j A_d[p_d][j] u(p | p_d = j) + b_d[p_d] f_d(p) ]
// with f_d the inflow neighbour's face value (A_d, b_d carry the speed and
// the upwind side, `hd_tables.cpp`).  Work mapping as in hd_vpwave.hip: a
// wave holds CPW = 64 / N^2 consecutive cells, N^2 lanes per cell; a lane is
// an x-plane (the N^2 values of one v node) for the x0 / x1 terms and a
// v-plane (the N^2 values of one x node) for the v0 / v1 terms, and the
// wave changes layout through its own LDS (no workgroup barrier).
//   stage   u (coalesced global loads) -> LDS -> x-planes
//   X       P = sum_{d = x0, x1} A_d u + b_d f_d              x-layout
//   V       K = P + sum_{d = v0, v1} A_d u + b_d f_d          v-layout
//   out     coalesced stores (LSRK: sol += b dt K, ti = sol + a dt K)
// The x faces are lines of the x-planes of the neighbour and the v faces
// lines of its v-planes; a lane loads its own (16-byte-aligned runs across
// the lanes of a cell).  FP32 applies every 1D operator in difference form
// A (u - u_o) + b (f - u_o), u_o the line's outflow value (A 1 + b = 0): no
// FP32 product carries the magnitude of u, and the face terms telescope.
#include "hd_device.h"
#include "hd_wave_common.h"

#include <algorithm>

namespace hd
{
namespace cvw
{
using wave::along0;
using wave::along0_acc;
using wave::along1;
using wave::along1_acc;
using wave::cptr;
using wave::launder;
using wave::nbr;
using wave::wave_sync;

template <int N>
struct Geo
{
  static constexpr int N2 = N * N, N4 = N2 * N2;
  static constexpr int CPW = 64 / N2;             // cells per wave
  static constexpr int LANES = CPW * N2;          // busy lanes
  static constexpr int P = N2 | 1;                // LDS row stride (odd)
  static constexpr int SLOT = 2 * N2 * P;         // LDS elements per cell
  static constexpr int KU = (CPW * N4 + 63) / 64; // coalesced values/lane
};

// One wave's work on one operator: groups g_first, g_first + g_stride, ...
// below g_end (group g: cells g CPW, ..., g CPW + CPW - 1), with the LDS of
// CPW cells.
template <int N, typename T, int MODE>
__device__ __forceinline__ void
cv_block(T *lds, const T *__restrict__ src, T *__restrict__ dst,
         const T *__restrict__ coef, const MeshDev &mesh,
         const InflowDev &inflow, const StageDev &stage, long long g_first,
         long long g_end, long long g_stride)
{
  using G = Geo<N>;
  constexpr int N2 = G::N2, N4 = G::N4, CPW = G::CPW, P = G::P;
  constexpr int KU = G::KU, SLOT = G::SLOT;
  const cptr<T> cc0 = (cptr<T>)coef;
  cptr<T> cc = cc0;
  // A_d (N x N, row = output node) and b_d (N) per direction
  auto A = [&](int d, int r) -> T { return cc[d * (N2 + N) + r]; };
  auto Bv = [&](int d, int i) -> T { return cc[d * (N2 + N) + N2 + i]; };
  const int n0 = mesh.n[0], n1 = mesh.n[1], n2 = mesh.n[2], n3 = mesh.n[3];
  const long long s2 = (long long)n0 * n1, s3 = s2 * n2;
  const long long n_cells = mesh.n_cells;
  // inflow side per direction: 0 left neighbour (its node N-1), 1 right
  // neighbour (its node 0), -1 no transport
  int side[4];
#pragma unroll
  for (int d = 0; d < 4; ++d)
    side[d] = inflow.side[d] > 0 ? 0 : (inflow.side[d] < 0 ? 1 : -1);

  int lane, s, t;
  bool busy;
  T *B0, *B1;
  // the lane and what derives from it are made opaque once per group (see
  // hd_vpwave.hip: no lane-dependent addresses held across the loop)
  auto lane_setup = [&]() {
    lane = (int)threadIdx.x;
    asm volatile("" : "+v"(lane));
    busy = G::LANES == 64 || lane < G::LANES;
    s = busy ? lane / N2 : 0;
    t = lane % N2;
    B0 = lds + s * SLOT;
    B1 = B0 + N2 * P;
  };
  lane_setup();
  auto lofs = [&](int k) -> int {
    if constexpr (64 % N2 == 0 && N4 % 64 == 0)
      {
        constexpr int M = N4 / 64;
        return (k / M) * SLOT + (64 / N2) * (k % M) * P + (lane / N2) * P +
               lane % N2;
      }
    else if constexpr (64 % N4 == 0)
      return (lane / N4 + (64 / N4) * k) * SLOT + (lane % N4) / N2 * P +
             lane % N2;
    else
      {
        const int e = lane + 64 * k, ss = e / N4, tt = e % N4;
        return ss * SLOT + (tt / N2) * P + tt % N2;
      }
  };
  auto for_values = [&](int n_in, auto f) {
    if (n_in == CPW)
      {
#pragma unroll
        for (int k = 0; k < KU; ++k)
          {
            const int e = lane + 64 * k;
            if (64 * KU == CPW * N4 || e < CPW * N4)
              f(k, e);
          }
      }
    else
      {
#pragma unroll
        for (int k = 0; k < KU; ++k)
          {
            const int e = lane + 64 * k;
            if (e < n_in * N4)
              f(k, e);
          }
      }
  };
  // cell coordinates and the inflow neighbour across direction d
  auto neighbour = [&](long long c, int d) -> long long {
    const int c0 = (int)(c % n0), c1 = (int)((c / n0) % n1);
    const int c2 = (int)((c / s2) % n2), c3 = (int)(c / s3);
    const int sd = side[d];
    return d == 0   ? nbr(c, c0, n0, 1, sd)
           : d == 1 ? nbr(c, c1, n1, n0, sd)
           : d == 2 ? nbr(c, c2, n2, s2, sd)
                    : nbr(c, c3, n3, s3, sd);
  };

  // the next group's values and all four inflow faces of this lane's cell
  // travel into registers while the current group computes: x faces as
  // lines of this lane's x-plane (x-layout, v node t), v faces as lines of
  // its v-plane (v-layout, x node t)
  T pu[KU], pf[4][N];
  auto prefetch = [&](long long g) {
    const long long c0 = g * CPW;
    const int n_in = (int)std::min<long long>(CPW, n_cells - c0);
    const long long base = c0 * N4;
    for_values(n_in, [&](int k, int e) { pu[k] = src[base + e]; });
    const bool act = busy && s < n_in;
    const long long c = c0 + s;
#pragma unroll
    for (int d = 0; d < 4; ++d)
      {
        const long long nb = act && side[d] >= 0 ? neighbour(c, d) : 0;
        const int fn = side[d] == 0 ? N - 1 : 0;
#pragma unroll
        for (int m = 0; m < N; ++m)
          {
            // node index of face value m: x faces vary the other x digit
            // at v node t, v faces the other v digit at x node t
            const int idx = d == 0   ? fn + N * m + N2 * t
                            : d == 1 ? m + N * fn + N2 * t
                            : d == 2 ? t + N2 * (fn + N * m)
                                     : t + N2 * (m + N * fn);
            pf[d][m] = act && side[d] >= 0 ? src[nb * N4 + idx] : (T)0;
          }
      }
  };

  long long g = g_first;
  if (g < g_end)
    prefetch(g);
  for (; g < g_end; g += g_stride)
    {
      cc = launder(cc0);
      lane_setup();
      const long long c0 = g * CPW;
      const int n_in = (int)std::min<long long>(CPW, n_cells - c0);
      const long long base = c0 * N4;
      T fx[4][N];
#pragma unroll
      for (int d = 0; d < 4; ++d)
#pragma unroll
        for (int m = 0; m < N; ++m)
          fx[d][m] = pf[d][m];
      // stage: the cells into B1 (rows of stride P per v node), then this
      // lane's x-plane into registers
      wave_sync();
      for_values(n_in, [&](int k, int) { lds[lofs(k) + N2 * P] = pu[k]; });
      wave_sync();
      // the next group's loads go out now
      if (g + g_stride < g_end)
        prefetch(g + g_stride);

      // X: P = sum over x0, x1 of A (u - u_o) + b (f - u_o) (FP32) or
      // A u + b f (FP64), along the lines of this lane's x-plane
      if (busy)
        {
          T u[N2], y[N2];
#pragma unroll
          for (int i = 0; i < N2; ++i)
            {
              u[i] = B1[t * P + i];
              y[i] = 0;
            }
#pragma unroll
          for (int d = 0; d < 2; ++d)
            {
              if (side[d] < 0)
                continue;
              const bool right = side[d] == 1; // outflow node 0
              T dd[N2], uo[N];
#pragma unroll
              for (int m = 0; m < N; ++m)
                {
                  // line m: along x0 at x1 = m (d = 0), along x1 at x0 = m
                  uo[m] = d == 0 ? (right ? u[N * m] : u[N - 1 + N * m])
                                 : (right ? u[m] : u[m + N * (N - 1)]);
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    {
                      const int j = d == 0 ? k + N * m : m + N * k;
                      dd[j] = sizeof(T) == 4 ? u[j] - uo[m] : u[j];
                    }
                }
              if (d == 0)
                along0_acc<N>(dd, y, [&](int r) { return A(0, r); });
              else
                along1_acc<N>(dd, y, [&](int r) { return A(1, r); });
#pragma unroll
              for (int m = 0; m < N; ++m)
                {
                  const T df = sizeof(T) == 4 ? fx[d][m] - uo[m] : fx[d][m];
#pragma unroll
                  for (int i = 0; i < N; ++i)
                    y[d == 0 ? i + N * m : m + N * i] += Bv(d, i) * df;
                }
            }
#pragma unroll
          for (int i = 0; i < N2; ++i)
            B0[t * P + i] = y[i];
        }
      wave_sync();
      // the solution (LSRK stage), used at the end
      double ps[KU];
      if constexpr (MODE == 1)
        for_values(n_in, [&](int k, int e) { ps[k] = stage.sol[base + e]; });
      // V: K = P + sum over v0, v1, along the lines of this lane's v-plane
      if (busy)
        {
          T u[N2], y[N2];
#pragma unroll
          for (int j = 0; j < N2; ++j)
            {
              u[j] = B1[j * P + t];
              y[j] = B0[j * P + t];
            }
#pragma unroll
          for (int e = 0; e < 2; ++e)
            {
              const int d = 2 + e;
              if (side[d] < 0)
                continue;
              const bool right = side[d] == 1;
              T dd[N2], uo[N];
#pragma unroll
              for (int m = 0; m < N; ++m)
                {
                  uo[m] = e == 0 ? (right ? u[N * m] : u[N - 1 + N * m])
                                 : (right ? u[m] : u[m + N * (N - 1)]);
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    {
                      const int j = e == 0 ? k + N * m : m + N * k;
                      dd[j] = sizeof(T) == 4 ? u[j] - uo[m] : u[j];
                    }
                }
              if (e == 0)
                along0_acc<N>(dd, y, [&](int r) { return A(2, r); });
              else
                along1_acc<N>(dd, y, [&](int r) { return A(3, r); });
#pragma unroll
              for (int m = 0; m < N; ++m)
                {
                  const T df = sizeof(T) == 4 ? fx[d][m] - uo[m] : fx[d][m];
#pragma unroll
                  for (int i = 0; i < N; ++i)
                    y[e == 0 ? i + N * m : m + N * i] += Bv(d, i) * df;
                }
            }
#pragma unroll
          for (int j = 0; j < N2; ++j)
            B0[j * P + t] = y[j];
        }
      wave_sync();
      // out: coalesced
      for_values(n_in, [&](int k, int e) {
        const T a = lds[lofs(k)];
        if constexpr (MODE == 0)
          dst[base + e] = a;
        else
          {
            const double so = ps[k];
            stage.sol[base + e] = so + stage.b_dt * (double)a;
            if (stage.ti_out)
              static_cast<T *>(stage.ti_out)[base + e] =
                (T)(so + stage.a_dt * (double)a);
          }
      });
    }
}

// single operator: persistent waves, each XCD walks a contiguous range of
// groups
template <int N, typename T, int MODE>
__global__ __launch_bounds__(64) void cv_kernel(
  const T *__restrict__ src, T *__restrict__ dst, const T *__restrict__ coef,
  MeshDev mesh, InflowDev inflow, StageDev stage, long long n_groups)
{
  using G = Geo<N>;
  __shared__ __align__(16) T lds[G::CPW * G::SLOT];
  const long long bpx = gridDim.x / 8;
  const long long per = (n_groups + 7) / 8;
  const long long xcd = blockIdx.x % 8;
  cv_block<N, T, MODE>(lds, src, dst, coef, mesh, inflow, stage,
                       xcd * per + blockIdx.x / 8,
                       std::min(n_groups, (xcd + 1) * per), bpx);
}

template <int N, typename T, int MODE>
int waves_per_cu()
{
  static int w = 0;
  if (w == 0)
    {
      int b = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b, cv_kernel<N, T, MODE>, 64, 0) != hipSuccess ||
          b < 1)
        b = 1;
      w = b;
    }
  return w;
}

template <int N, typename T>
hipError_t launch(int mode, const T *src, T *dst, const T *coef,
                  const MeshDev &mesh, const InflowDev &inflow,
                  const StageDev &stage, hipStream_t stream)
{
  using G = Geo<N>;
  const long long n_groups = (mesh.n_cells + G::CPW - 1) / G::CPW;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess ||
      cus < 8)
    cus = 256;
  const int wpc =
    mode == 0 ? waves_per_cu<N, T, 0>() : waves_per_cu<N, T, 1>();
  long long grid = (long long)cus * wpc;
  grid = std::min<long long>(grid, (n_groups + 7) / 8 * 8);
  grid = std::max<long long>(8, grid / 8 * 8);
  if (mode == 0)
    hipLaunchKernelGGL((cv_kernel<N, T, 0>), dim3((unsigned)grid), dim3(64),
                       0, stream, src, dst, coef, mesh, inflow, stage,
                       n_groups);
  else
    hipLaunchKernelGGL((cv_kernel<N, T, 1>), dim3((unsigned)grid), dim3(64),
                       0, stream, src, dst, coef, mesh, inflow, stage,
                       n_groups);
  return hipGetLastError();
}

template <typename T>
hipError_t dispatch(int N, int mode, const T *src, T *dst, const T *coef,
                    const MeshDev &mesh, const InflowDev &inflow,
                    const StageDev &stage, hipStream_t stream)
{
  switch (N)
    {
      case 2:
        return launch<2, T>(mode, src, dst, coef, mesh, inflow, stage,
                            stream);
      case 3:
        return launch<3, T>(mode, src, dst, coef, mesh, inflow, stage,
                            stream);
      case 5:
        if constexpr (sizeof(T) == 4)
          return launch<5, T>(mode, src, dst, coef, mesh, inflow, stage,
                              stream);
        else
          return hipErrorInvalidValue;
      default:
        return hipErrorInvalidValue;
    }
}
} // namespace cvw

namespace cvw
{
// a batch of operators of degree N (the batch work split of hd_batch.hip:
// tasks of logical units = groups of CPW cells); one wave per block, the
// same block code as the single-operator kernel, so a batched apply or
// stage equals the single-operator call bit for bit
template <int N, typename T, int MODE>
__global__ __launch_bounds__(64) void batch_kernel(
  const BatchOp *__restrict__ ops, const BatchTask *__restrict__ tasks,
  BatchPlan plan, BatchPtrs ptrs)
{
  using G = Geo<N>;
  __shared__ __align__(16) T lds[G::CPW * G::SLOT];
  const int b = (blockIdx.x % 8) * plan.per + blockIdx.x / 8;
  const int t_end = plan.block_start[b + 1];
  for (int tt = plan.block_start[b]; tt < t_end; ++tt)
    {
      const BatchTask tk = tasks[tt];
      const BatchOp &o = ops[tk.op];
      MeshDev mesh;
      InflowDev inflow;
      mesh.dim = o.mesh.dim;
      mesh.n_cells = o.mesh.n_cells;
      mesh.inner = o.mesh.inner;
#pragma unroll
      for (int d = 0; d < kMaxDim; ++d)
        {
          mesh.n[d] = o.mesh.n[d];
          mesh.stride[d] = o.mesh.stride[d];
          mesh.lo[d] = o.mesh.lo[d];
          mesh.h[d] = o.mesh.h[d];
          inflow.side[d] = o.inflow.side[d];
        }
      StageDev stage{};
      stage.a_dt = ptrs.a_dt;
      stage.b_dt = ptrs.b_dt;
      stage.sol = MODE == 0 ? nullptr : ptrs.sol[tk.op];
      stage.ti_out = MODE == 1 ? ptrs.ti[tk.op] : nullptr;
      cv_block<N, T, MODE == 0 ? 0 : 1>(
        lds, static_cast<const T *>(ptrs.src[tk.op]),
        MODE == 0 ? static_cast<T *>(ptrs.dst[tk.op]) : nullptr,
        static_cast<const T *>(o.coef), mesh, inflow, stage, tk.lb, tk.le,
        tk.stride);
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
    hipLaunchKernelGGL(kernel, dim3(8 * plan.per), dim3(64), 0, stream, ops,
                       tasks, plan, ptrs);
  };
  if (mode == 0)
    go(batch_kernel<N, T, 0>);
  else if (mode == 1)
    go(batch_kernel<N, T, 1>);
  else
    go(batch_kernel<N, T, 2>);
  return hipGetLastError();
}
} // namespace cvw

// Which 2+2 operators run here, measured on MI355X (GDoF/s, 2+2, about
// 10^8 DoF, this kernel against the tile kernel of hd_collapsed.hip):
//   k = 2 (N = 3): FP64 stage 114.9 vs 84.7, FP32 stage 154.7 vs 101.6,
//                  FP32 apply 292.7 vs 280.4, FP64 apply 171.6 vs 171.5
//   k = 4 (N = 5), FP32: stage 178.3 vs 140.6, apply 312.4 vs 373.4 (the
//                  stage is what an LSRK run repeats); FP64 stage 109.6 vs
//                  113.9 and apply 184.5 vs 195.9 stay on the tile kernel
//   k = 3: within 5 %; k = 5: slower throughout (a wave holds one cell on
//   36 of its 64 lanes)
bool collapsed_wave_supported(int D, int N, int precision)
{
  return D == 4 && (N == 2 || N == 3 || (N == 5 && precision != 0));
}

int collapsed_wave_cells_per_unit(int N)
{
  return N == 2 ? cvw::Geo<2>::CPW
                : (N == 3 ? cvw::Geo<3>::CPW : cvw::Geo<5>::CPW);
}

int collapsed_wave_blocks_per_cu(int N, int precision)
{
  if (N == 2)
    return precision == 0 ? cvw::waves_per_cu<2, double, 1>()
                          : cvw::waves_per_cu<2, float, 1>();
  if (N == 3)
    return precision == 0 ? cvw::waves_per_cu<3, double, 1>()
                          : cvw::waves_per_cu<3, float, 1>();
  return cvw::waves_per_cu<5, float, 1>();
}

hipError_t collapsed_wave_batch(int N, int precision, int mode,
                                const BatchOp *ops, const BatchTask *tasks,
                                const BatchPlan &plan, const BatchPtrs &ptrs,
                                hipStream_t stream)
{
  if (N == 2)
    return precision == 0
             ? cvw::batch_launch<2, double>(mode, ops, tasks, plan, ptrs,
                                            stream)
             : cvw::batch_launch<2, float>(mode, ops, tasks, plan, ptrs,
                                           stream);
  if (N == 3)
    return precision == 0
             ? cvw::batch_launch<3, double>(mode, ops, tasks, plan, ptrs,
                                            stream)
             : cvw::batch_launch<3, float>(mode, ops, tasks, plan, ptrs,
                                           stream);
  if (N == 5 && precision != 0)
    return cvw::batch_launch<5, float>(mode, ops, tasks, plan, ptrs, stream);
  return hipErrorInvalidValue;
}

hipError_t collapsed_wave_f64(int N, int mode, const double *src, double *dst,
                              const double *coef, const MeshDev &mesh,
                              const InflowDev &inflow, const StageDev &stage,
                              hipStream_t stream)
{
  return cvw::dispatch<double>(N, mode, src, dst, coef, mesh, inflow, stage,
                               stream);
}

hipError_t collapsed_wave_f32(int N, int mode, const float *src, float *dst,
                              const float *coef, const MeshDev &mesh,
                              const InflowDev &inflow, const StageDev &stage,
                              hipStream_t stream)
{
  return cvw::dispatch<float>(N, mode, src, dst, coef, mesh, inflow, stage,
                              stream);
}
} // namespace hd
