// hd_wave_common.h -- helpers of the wave-per-group kernels (hd_vpwave.hip,
// hd_cvwave.hip): one wave owns a group of cells, a lane holds an N x N
// plane of one cell in registers, and the wave changes between plane
// layouts through its own LDS.
#pragma once

#include <hip/hip_runtime.h>

namespace hd
{
namespace wave
{
template <class T>
using cptr = const __attribute__((address_space(4))) T *;

// orders the LDS accesses of one wave (lanes of one wave see each other's
// LDS writes in program order; this keeps the compiler from moving them)
__device__ inline void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// the same address, opaque to the compiler: the coefficient tables are
// re-read (scalar loads) in every group instead of being hoisted out of the
// group loop into a register-hungry copy
template <class T>
__device__ inline cptr<T> launder(cptr<T> p)
{
  unsigned long long v = (unsigned long long)p;
  asm volatile("" : "+s"(v));
  return (cptr<T>)v;
}

// periodic neighbour of cell c across a direction with n cells, position
// ci and cell-id stride st; side 0: left, 1: right
__device__ inline long long nbr(long long c, int ci, int n, long long st,
                                int side)
{
  if (side == 0)
    return ci == 0 ? c + (long long)(n - 1) * st : c - st;
  return ci == n - 1 ? c - (long long)(n - 1) * st : c + st;
}

// y = M along the first digit of an N x N plane (in[a + N b] -> y[i + N b])
template <int N, typename T, typename F>
__device__ inline void along0(const T *in, T *y, F coef)
{
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int i = 0; i < N; ++i)
      {
        T a = 0;
#pragma unroll
        for (int k = 0; k < N; ++k)
          a += coef(i * N + k) * in[k + N * b];
        y[i + N * b] = a;
      }
}

// y = M along the second digit (in[a + N b] -> y[a + N i])
template <int N, typename T, typename F>
__device__ inline void along1(const T *in, T *y, F coef)
{
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int a = 0; a < N; ++a)
      {
        T s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k)
          s += coef(i * N + k) * in[a + N * k];
        y[a + N * i] = s;
      }
}

// y += M along the first / second digit
template <int N, typename T, typename F>
__device__ inline void along0_acc(const T *in, T *y, F coef)
{
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int k = 0; k < N; ++k)
        y[i + N * b] += coef(i * N + k) * in[k + N * b];
}
template <int N, typename T, typename F>
__device__ inline void along1_acc(const T *in, T *y, F coef)
{
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int k = 0; k < N; ++k)
        y[a + N * i] += coef(i * N + k) * in[a + N * k];
}

} // namespace wave
} // namespace hd
