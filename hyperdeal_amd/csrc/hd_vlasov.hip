// hd_vlasov.hip -- the Vlasov-Poisson pieces around the advection apply
// (examples/vlasov_poisson/include/application.h:516-602), gfx950:
//   rho(x) = int f dv          vector_tools.h:230-315 (Gauss-Lobatto v rule
//                              on the nodal values)
//   b = -M (rho - mean), b -= mean                      application.h:531-568
//   A phi = b                  periodic SIPG -Laplacian, poisson.h:100-254,
//                              solved with FP64 conjugate gradients from the
//                              previous potential
//   E = grad phi at the x quadrature points   derivative_container.h:158-251
// and the phase-space diagnostics (diagnostics.h:28-141).
//
// The x-space problem is small (32^2 cells, 16,384 DoF at k=3), so the whole
// solve runs in one workgroup: every thread owns whole x cells, the search
// direction lives in LDS (the SIPG face terms read the neighbour cells from
// there), dot products are block reductions, and the iteration count stays
// on the device -- no host round trip inside a Runge-Kutta stage.
#include "hd_device.h"
#include "hd_vp_table.h"

#include <algorithm>

namespace hd
{
namespace vlasov
{
constexpr int kSolverThreads = 1024;

// ---- rho: partial sums over chunks of v cells, then a fixed-order sum ----
// copy32 / copy64 (may be null): the values read are also written there
// (the first LSRK stage's input register, from the FP64 solution)
__global__ __launch_bounds__(256) void density_partial(
  const double *__restrict__ f, const float *__restrict__ f32,
  VpDev vp, long long n_x_cells, long long n_v_cells, int chunk,
  double *__restrict__ partial, float *__restrict__ copy32,
  double *__restrict__ copy64)
{
  // block: one x cell, v cells [blockIdx.y * chunk, ...)
  const long long xc = blockIdx.x;
  const long long v0 = (long long)blockIdx.y * chunk;
  const long long v1 = min(v0 + chunk, n_v_cells);
  const int dcx = vp.dofs_x, dcv = vp.dofs_v, dc = dcx * dcv;
  __shared__ double red[256];
  // thread t accumulates value (x node i, v node j) for i = t % dcx
  double acc = 0.0;
  for (long long vc = v0; vc < v1; ++vc)
    {
      const long long cell = xc + n_x_cells * vc;
      for (int t = threadIdx.x; t < dc; t += blockDim.x)
        {
          const double v = f ? f[cell * dc + t] : (double)f32[cell * dc + t];
          if (copy32)
            copy32[cell * dc + t] = (float)v;
          if (copy64)
            copy64[cell * dc + t] = v;
          acc += v * vp.wv[t / dcx];
        }
    }
  // the block size is a multiple of dcx: thread t only saw x node t % dcx
  red[threadIdx.x] = acc;
  __syncthreads();
  if ((int)threadIdx.x < dcx)
    {
      double s = 0.0;
      for (int t = threadIdx.x; t < (int)blockDim.x; t += dcx)
        s += red[t];
      partial[(xc * gridDim.y + blockIdx.y) * dcx + threadIdx.x] = s;
    }
}

__global__ void density_reduce(const double *__restrict__ partial,
                               int n_chunks, int dcx, long long n_x_dofs,
                               double *__restrict__ rho)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_x_dofs)
    return;
  const long long xc = i / dcx;
  const int node = (int)(i - xc * dcx);
  double s = 0.0;
  for (int c = 0; c < n_chunks; ++c)
    s += partial[(xc * n_chunks + c) * dcx + node];
  rho[i] = s;
}

// ---- the single-block x-space solver ----
template <int DX, int N>
struct XCell
{
  static constexpr int DOFS = ipow(N, DX);
};

__device__ inline double block_sum(double v, double *red)
{
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s /= 2)
    {
      if ((int)threadIdx.x < s)
        red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
  const double r = red[0];
  __syncthreads();
  return r;
}

// y = (A_0 (x) ... (x) A_{DX-1}) x for N x N matrices, in registers
template <int DX, int N>
__device__ inline void tensor_nn(const double *const *A, const double *x,
                                 double *y)
{
  constexpr int DOFS = ipow(N, DX);
  double cur[DOFS], nxt[DOFS];
#pragma unroll
  for (int i = 0; i < DOFS; ++i)
    cur[i] = x[i];
#pragma unroll
  for (int d = 0; d < DX; ++d)
    {
      const int sd = ipow(N, d);
#pragma unroll
      for (int i = 0; i < DOFS; ++i)
        {
          const int id = (i / sd) % N, rest = i - id * sd;
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < N; ++j)
            s += A[d][id * N + j] * cur[rest + j * sd];
          nxt[i] = s;
        }
#pragma unroll
      for (int i = 0; i < DOFS; ++i)
        cur[i] = nxt[i];
    }
#pragma unroll
  for (int i = 0; i < DOFS; ++i)
    y[i] = cur[i];
}

// periodic SIPG apply for the cell owned by this thread; src lives in LDS
// (mean already removed by the caller)
template <int DX, int N>
__device__ inline void sipg_cell(const XDev &xd, const double *src,
                                 long long xc, double *y)
{
  constexpr int DOFS = ipow(N, DX);
  const double *x = src + xc * DOFS;
  int idx[3] = {0, 0, 0};
  {
    long long r = xc;
#pragma unroll
    for (int d = 0; d < DX; ++d)
      {
        idx[d] = (int)(r % xd.n[d]);
        r /= xd.n[d];
      }
  }
#pragma unroll
  for (int i = 0; i < DOFS; ++i)
    y[i] = 0.0;
  // cell: sum_d (x)_{j != d} h_j M (x) K / h_d
#pragma unroll
  for (int d = 0; d < DX; ++d)
    {
      double Ms[3][N * N], t[DOFS];
      const double *A[3];
#pragma unroll
      for (int j = 0; j < DX; ++j)
        {
#pragma unroll
          for (int e = 0; e < N * N; ++e)
            Ms[j][e] = j == d ? xd.K[e] / xd.h[j] : xd.M[e] * xd.h[j];
          A[j] = Ms[j];
        }
      tensor_nn<DX, N>(A, x, t);
#pragma unroll
      for (int i = 0; i < DOFS; ++i)
        y[i] += t[i];
    }
  // faces: both faces of this cell in every direction; this cell is the
  // minus side of its right face and the plus side of its left face
  const double kk = xd.kk;
#pragma unroll
  for (int d = 0; d < DX; ++d)
    {
      const int sd = ipow(N, d);
      const double h = xd.h[d];
      const double sigma = kk * (2.0 / h);
      long long stride = 1;
      for (int j = 0; j < d; ++j)
        stride *= xd.n[j];
#pragma unroll
      for (int side = 0; side < 2; ++side)
        {
          // neighbour across this face
          const long long nb =
            side ? (idx[d] == xd.n[d] - 1 ? xc - (long long)(xd.n[d] - 1) * stride
                                          : xc + stride)
                 : (idx[d] == 0 ? xc + (long long)(xd.n[d] - 1) * stride
                                : xc - stride);
          const double *xn = src + nb * DOFS;
          // 1D quantities per tangential node, tested in this cell only
          double fa[DOFS];
#pragma unroll
          for (int i = 0; i < DOFS; ++i)
            fa[i] = 0.0;
#pragma unroll
          for (int tf = 0; tf < DOFS / N; ++tf)
            {
              // base index of tangential node tf (digit d = 0)
              int base = 0, rem = tf, pw = 1;
#pragma unroll
              for (int j = 0; j < DX; ++j)
                {
                  if (j != d)
                    {
                      base += (rem % N) * pw;
                      rem /= N;
                    }
                  pw *= N;
                }
              // minus = left cell of the face, plus = right cell
              const double *xl = side ? x + base : xn + base;
              const double *xr = side ? xn + base : x + base;
              const double um = xl[(N - 1) * sd], upl = xr[0];
              double dm = 0.0, dp = 0.0;
#pragma unroll
              for (int i = 0; i < N; ++i)
                {
                  dm += xd.d1[i] * xl[i * sd] / h;
                  dp += xd.d0[i] * xr[i * sd] / h;
                }
              const double jump = um - upl, avgd = 0.5 * (dm + dp);
              const double val = sigma * jump - avgd;
#pragma unroll
              for (int i = 0; i < N; ++i)
                {
                  if (side) // this cell is the minus side
                    fa[base + i * sd] = (i == N - 1 ? val : 0.0) -
                                        0.5 * jump * xd.d1[i] / h;
                  else // plus side
                    fa[base + i * sd] = (i == 0 ? -val : 0.0) -
                                        0.5 * jump * xd.d0[i] / h;
                }
            }
          // tangential mass
          double Ms[3][N * N], t[DOFS];
          const double *A[3];
#pragma unroll
          for (int j = 0; j < DX; ++j)
            {
#pragma unroll
              for (int e = 0; e < N * N; ++e)
                Ms[j][e] = j == d ? (e / N == e % N ? 1.0 : 0.0)
                                  : xd.M[e] * xd.h[j];
              A[j] = Ms[j];
            }
          tensor_nn<DX, N>(A, fa, t);
#pragma unroll
          for (int i = 0; i < DOFS; ++i)
            y[i] += t[i];
        }
    }
}

// One block: RHS, CG from phi (kept between calls), grad phi.  The vectors a
// thread only reads and writes for its own cells (residual, iterate, A p,
// direction) live in global scratch; the matvec input, which the SIPG face
// terms read across cells, lives in LDS.
template <int DX, int N>
__global__ __launch_bounds__(kSolverThreads) void poisson_kernel(
  XDev xd, const double *__restrict__ rho_in, double *__restrict__ phi,
  const double *__restrict__ guess, double *__restrict__ efield,
  double *__restrict__ work, double rel_tol, int max_it,
  long long *__restrict__ stats)
{
  constexpr int DOFS = ipow(N, DX);
  extern __shared__ double sm[];
  const long long nc = xd.n_cells;
  const long long nx = nc * DOFS;
  double *p = sm;       // nx, the matvec input
  double *red = sm + nx; // blockDim
  double *r = work, *Ap = work + nx, *dir = work + 2 * nx;
  const int T = blockDim.x;

  // RHS: rho -= mean, b = -M rho, b -= mean (r holds b first)
  double s = 0.0;
  for (long long c = threadIdx.x; c < nc; c += T)
    for (int i = 0; i < DOFS; ++i)
      s += rho_in[c * DOFS + i];
  const double mean_rho = block_sum(s, red) / (double)nx;
  double Ms[3][N * N];
  const double *MA[3];
  for (int j = 0; j < DX; ++j)
    {
      for (int e = 0; e < N * N; ++e)
        Ms[j][e] = xd.M[e] * xd.h[j];
      MA[j] = Ms[j];
    }
  s = 0.0;
  for (long long c = threadIdx.x; c < nc; c += T)
    {
      double x[DOFS], y[DOFS];
      for (int i = 0; i < DOFS; ++i)
        x[i] = rho_in[c * DOFS + i] - mean_rho;
      tensor_nn<DX, N>(MA, x, y);
      for (int i = 0; i < DOFS; ++i)
        {
          r[c * DOFS + i] = -y[i];
          s += -y[i];
        }
    }
  const double mean_b = block_sum(s, red) / (double)nx;
  double bb = 0.0;
  for (long long c = threadIdx.x; c < nc; c += T)
    for (int i = 0; i < DOFS; ++i)
      {
        const double v = r[c * DOFS + i] - mean_b;
        r[c * DOFS + i] = v;
        bb += v * v;
      }
  bb = block_sum(bb, red);

  // out = A (in - mean(in)) with in in global memory
  auto apply_A = [&](const double *in, double *out) {
    double sm_ = 0.0;
    for (long long c = threadIdx.x; c < nc; c += T)
      for (int i = 0; i < DOFS; ++i)
        sm_ += in[c * DOFS + i];
    const double mean = block_sum(sm_, red) / (double)nx;
    for (long long c = threadIdx.x; c < nc; c += T)
      for (int i = 0; i < DOFS; ++i)
        p[c * DOFS + i] = in[c * DOFS + i] - mean;
    __syncthreads();
    for (long long c = threadIdx.x; c < nc; c += T)
      {
        double y[DOFS];
        sipg_cell<DX, N>(xd, p, c, y);
        for (int i = 0; i < DOFS; ++i)
          out[c * DOFS + i] = y[i];
      }
    __syncthreads();
  };
  // r = b - A phi
  apply_A(phi, Ap);
  double alpha0 = 1.0, beta0 = 0.0, rr_warm = -1.0;
  if (guess)
    {
      // start from the minimiser of the energy 1/2 x.A x - b.x over
      // x = alpha phi + beta g (phi the previous potential, g the guess):
      // [phi.A phi, phi.A g; phi.A g, g.A g] (alpha, beta) = (phi.b, g.b)
      apply_A(guess, dir); // dir = A g
      double d[6] = {0, 0, 0, 0, 0, 0};
      for (long long c = threadIdx.x; c < nc; c += T)
        for (int i = 0; i < DOFS; ++i)
          {
            const long long g = c * DOFS + i;
            const double b = r[g], x = phi[g], y = guess[g];
            d[0] += x * Ap[g];
            d[1] += x * dir[g];
            d[2] += y * dir[g];
            d[3] += x * b;
            d[4] += y * b;
            d[5] += (b - Ap[g]) * (b - Ap[g]);
          }
      for (int j = 0; j < 6; ++j)
        d[j] = block_sum(d[j], red);
      rr_warm = d[5];
      const double det = d[0] * d[2] - d[1] * d[1];
      if (d[0] > 0.0 && d[2] > 0.0 && det > 1e-12 * d[0] * d[2])
        {
          alpha0 = (d[3] * d[2] - d[4] * d[1]) / det;
          beta0 = (d[0] * d[4] - d[1] * d[3]) / det;
        }
      else if (!(d[0] > 0.0) && d[2] > 0.0)
        {
          alpha0 = 0.0;
          beta0 = d[4] / d[2];
        }
      if (!isfinite(alpha0) || !isfinite(beta0))
        {
          alpha0 = 1.0;
          beta0 = 0.0;
        }
    }
  double rr = 0.0;
  for (long long c = threadIdx.x; c < nc; c += T)
    for (int i = 0; i < DOFS; ++i)
      {
        const long long g = c * DOFS + i;
        if (guess)
          {
            phi[g] = alpha0 * phi[g] + beta0 * guess[g];
            Ap[g] = alpha0 * Ap[g] + beta0 * dir[g];
          }
        r[g] -= Ap[g];
        rr += r[g] * r[g];
      }
  rr = block_sum(rr, red);
  for (long long c = threadIdx.x; c < nc; c += T)
    for (int i = 0; i < DOFS; ++i)
      dir[c * DOFS + i] = r[c * DOFS + i];
  __syncthreads();
  const double rr_start = rr;
  const double stop = rel_tol * sqrt(bb);
  int it = 0;
  while (sqrt(rr) > stop && it < max_it)
    {
      apply_A(dir, Ap);
      double pAp = 0.0;
      for (long long c = threadIdx.x; c < nc; c += T)
        for (int i = 0; i < DOFS; ++i)
          pAp += dir[c * DOFS + i] * Ap[c * DOFS + i];
      pAp = block_sum(pAp, red);
      if (!(pAp > 0.0))
        break; // what is left of the residual lies in the null space
      const double alpha = rr / pAp;
      double rr2 = 0.0;
      for (long long c = threadIdx.x; c < nc; c += T)
        for (int i = 0; i < DOFS; ++i)
          {
            const long long g = c * DOFS + i;
            phi[g] += alpha * dir[g];
            r[g] -= alpha * Ap[g];
            rr2 += r[g] * r[g];
          }
      rr2 = block_sum(rr2, red);
      const double beta = rr2 / rr;
      rr = rr2;
      for (long long c = threadIdx.x; c < nc; c += T)
        for (int i = 0; i < DOFS; ++i)
          {
            const long long g = c * DOFS + i;
            dir[g] = r[g] + beta * dir[g];
          }
      ++it;
    }

  // grad phi at the x quadrature points
  for (long long c = threadIdx.x; c < nc; c += T)
    {
      double x[DOFS], g[DOFS];
      for (int i = 0; i < DOFS; ++i)
        x[i] = phi[c * DOFS + i];
      for (int e = 0; e < DX; ++e)
        {
          const double *A[3];
          for (int j = 0; j < DX; ++j)
            A[j] = j == e ? xd.G : xd.S;
          tensor_nn<DX, N>(A, x, g);
          for (int q = 0; q < DOFS; ++q)
            efield[(c * DOFS + q) * DX + e] = g[q] / xd.h[e];
        }
    }
  if (threadIdx.x == 0)
    {
      double *ds = reinterpret_cast<double *>(stats);
      const double nb = sqrt(bb) + 1e-300;
      stats[0] += it; // iterations, accumulated over solves
      stats[1] += 1;  // solves
      ds[2] = sqrt(rr) / nb;
      // relative residual of the start, and of the previous potential
      ds[3] = sqrt(rr_start) / nb;
      ds[4] = (guess ? sqrt(rr_warm) : sqrt(rr_start)) / nb;
      if (guess)
        {
          stats[5] += 1; // solves that started from a guess
          ds[6] = alpha0;
          ds[7] = beta0;
        }
    }
}

// ---- the 2D solver: tensor-product SIPG blocks, registers and LDS ----
// On a uniform periodic 2D mesh the SIPG matrix is A = L_0 (x) M_1 +
// M_0 (x) L_1 with L_d the 1D SIPG matrix of direction d (blocks coupling a
// cell to its left neighbour, itself, its right neighbour) and M_d the 1D
// mass matrix.  Table (8 N^2 doubles): L_0 left, self, right, L_1 left,
// self, right, M_0, M_1, each row-major (test node, trial node).
template <class T>
using cptr = const __attribute__((address_space(4))) T *;

// LDS stride of the transposed direction (one row per node): a compile-time
// constant, so every node's read is an immediate offset from one address
constexpr int kPStride = kSolverThreads;

// y = A p for the cell (cx, cy); p transposed in LDS: Pt[i * kPStride + cell]
template <int N>
__device__ inline void sipg2d(const double *Pt, int n0, int n1, int cx,
                              int cy, cptr<double> tab, double *y)
{
  constexpr int NN = N * N;
  // opaque to the optimiser: the table loads stay inside the CG loop as
  // scalar loads instead of being hoisted into vector registers
  asm volatile("" : "+s"(tab));
  const int cs = cx + n0 * cy;
  const int xn[3] = {(cx == 0 ? n0 - 1 : cx - 1) + n0 * cy, cs,
                     (cx == n0 - 1 ? 0 : cx + 1) + n0 * cy};
  const int yn[3] = {cx + n0 * (cy == 0 ? n1 - 1 : cy - 1), cs,
                     cx + n0 * (cy == n1 - 1 ? 0 : cy + 1)};
  double z[NN];
  // direction 0: z = sum_k (L_0k (x) I) p_k, y = (I (x) M_1) z
#pragma unroll
  for (int iy = 0; iy < N; ++iy)
    {
#pragma unroll
      for (int ix = 0; ix < N; ++ix)
        z[ix + N * iy] = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        {
          double q[N];
#pragma unroll
          for (int jx = 0; jx < N; ++jx)
            q[jx] = Pt[(jx + N * iy) * kPStride + xn[k]];
#pragma unroll
          for (int ix = 0; ix < N; ++ix)
#pragma unroll
            for (int jx = 0; jx < N; ++jx)
              z[ix + N * iy] += tab[k * NN + ix * N + jx] * q[jx];
        }
    }
#pragma unroll
  for (int iy = 0; iy < N; ++iy)
#pragma unroll
    for (int ix = 0; ix < N; ++ix)
      {
        double a = 0.0;
#pragma unroll
        for (int jy = 0; jy < N; ++jy)
          a += tab[7 * NN + iy * N + jy] * z[ix + N * jy];
        y[ix + N * iy] = a;
      }
  // direction 1: z = sum_k (I (x) L_1k) p_k, y += (M_0 (x) I) z
#pragma unroll
  for (int ix = 0; ix < N; ++ix)
    {
#pragma unroll
      for (int iy = 0; iy < N; ++iy)
        z[ix + N * iy] = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        {
          double q[N];
#pragma unroll
          for (int jy = 0; jy < N; ++jy)
            q[jy] = Pt[(ix + N * jy) * kPStride + yn[k]];
#pragma unroll
          for (int iy = 0; iy < N; ++iy)
#pragma unroll
            for (int jy = 0; jy < N; ++jy)
              z[ix + N * iy] += tab[(3 + k) * NN + iy * N + jy] * q[jy];
        }
    }
#pragma unroll
  for (int iy = 0; iy < N; ++iy)
#pragma unroll
    for (int ix = 0; ix < N; ++ix)
      {
        double a = 0.0;
#pragma unroll
        for (int jx = 0; jx < N; ++jx)
          a += tab[6 * NN + ix * N + jx] * z[jx + N * iy];
        y[ix + N * iy] += a;
      }
}

// sum over the block (W waves): wave shuffles, one barrier, every thread
// adds the W wave sums in the same order (uniform result).  Consecutive
// calls use alternate slots, so a slot is rewritten only after every thread
// has passed the barrier of the call in between.
template <int W, int slot>
__device__ inline double block_sum(double v, double *red)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
    v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0)
    red[slot * 16 + (threadIdx.x >> 6)] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < W; ++w)
    r += red[slot * 16 + w];
  return r;
}

// One block of 1024 / CPT threads; thread t owns the x cells t + k 1024/CPT,
// k < CPT (n_cells <= 1024): residuals and A p in registers, the search
// direction in LDS (transposed, so that a wave reads consecutive addresses),
// the iterate in global scratch (transposed).  Two cells per thread leave
// 256 registers per thread (no spills) and two independent matrix-vector
// products in flight.  Same algorithm and outputs as poisson_kernel.
template <int N, int CPT>
__global__ __launch_bounds__(kSolverThreads / CPT) void poisson2d_kernel(
  XDev xd, const double *__restrict__ tab_in,
  const double *__restrict__ rho_in, double *__restrict__ phi,
  const double *__restrict__ guess, double *__restrict__ efield,
  double *__restrict__ work, double rel_tol, int max_it,
  long long *__restrict__ stats, float *__restrict__ res_image,
  double *__restrict__ res_scale, int guess_alone)
{
  constexpr int NN = N * N, T = kSolverThreads / CPT, W = T / 64;
  extern __shared__ double sm[];
  const int nc = (int)xd.n_cells, n0 = xd.n[0], n1 = xd.n[1];
  const double nx = (double)nc * NN;
  double *Pt = sm;                  // NN x kPStride
  double *red = sm + NN * kPStride; // 2 x 16
  double *phit = work;              // NN x nc
  const cptr<double> tab = (cptr<double>)tab_in;
  const int t = threadIdx.x;
  bool own[CPT];
  int cl[CPT], cx[CPT], cy[CPT];
#pragma unroll
  for (int k = 0; k < CPT; ++k)
    {
      const int c = t + k * T;
      own[k] = c < nc;
      cl[k] = own[k] ? c : 0;
      cx[k] = cl[k] % n0;
      cy[k] = cl[k] / n0;
    }

  // b = -(M_0 (x) M_1)(rho - mean rho), then b -= mean b
  double r[CPT][NN], y[CPT][NN];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k)
#pragma unroll
    for (int i = 0; i < NN; ++i)
      {
        r[k][i] = own[k] ? rho_in[(long long)cl[k] * NN + i] : 0.0;
        s += r[k][i];
      }
  const double mean_rho = block_sum<W, 0>(s, red) / nx;
  s = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k)
    {
#pragma unroll
      for (int iy = 0; iy < N; ++iy)
#pragma unroll
        for (int ix = 0; ix < N; ++ix)
          {
            double a = 0.0;
#pragma unroll
            for (int jx = 0; jx < N; ++jx)
              a += tab[6 * NN + ix * N + jx] * (r[k][jx + N * iy] - mean_rho);
            y[k][ix + N * iy] = a;
          }
#pragma unroll
      for (int iy = 0; iy < N; ++iy)
#pragma unroll
        for (int ix = 0; ix < N; ++ix)
          {
            double a = 0.0;
#pragma unroll
            for (int jy = 0; jy < N; ++jy)
              a += tab[7 * NN + iy * N + jy] * y[k][ix + N * jy];
            r[k][ix + N * iy] = own[k] ? -a : 0.0;
            s += r[k][ix + N * iy];
          }
    }
  const double mean_b = block_sum<W, 1>(s, red) / nx;
  double bb = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k)
#pragma unroll
    for (int i = 0; i < NN; ++i)
      {
        r[k][i] = own[k] ? r[k][i] - mean_b : 0.0; // r holds b
        bb += r[k][i] * r[k][i];
      }
  bb = block_sum<W, 0>(bb, red);

  // the start: r = b - A phi (or the combination with the guess); phi goes
  // to the transposed iterate right away
#pragma unroll
  for (int k = 0; k < CPT; ++k)
    if (own[k])
#pragma unroll
      for (int i = 0; i < NN; ++i)
        {
          const double v = phi[(long long)cl[k] * NN + i];
          Pt[i * kPStride + cl[k]] = v;
          phit[(long long)i * nc + cl[k]] = v;
        }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < CPT; ++k)
    {
      if (own[k])
        sipg2d<N>(Pt, n0, n1, cx[k], cy[k], tab, y[k]);
      else
#pragma unroll
        for (int i = 0; i < NN; ++i)
          y[k][i] = 0.0;
    }
  double alpha0 = 1.0, beta0 = 0.0, rr_warm = -1.0;
  if (guess)
    {
      // the guess without its nodal mean: A g is the same, and the start
      // keeps the zero nodal mean of the iterates (Krylov updates lie in the
      // range of A, orthogonal to the constants)
      double sg = 0.0;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
#pragma unroll
          for (int i = 0; i < NN; ++i)
            sg += guess[(long long)cl[k] * NN + i];
      const double mean_g = block_sum<W, 1>(sg, red) / nx;
      double d[6] = {0, 0, 0, 0, 0, 0};
      __syncthreads(); // every thread has read Pt
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
#pragma unroll
          for (int i = 0; i < NN; ++i)
            Pt[i * kPStride + cl[k]] =
              guess[(long long)cl[k] * NN + i] - mean_g;
      __syncthreads();
      if (guess_alone)
        {
          // an exact guess: the start is the guess itself (a combination
          // with a previous phi that is nearly parallel to it would be an
          // ill-conditioned 2 x 2 system)
#pragma unroll
          for (int k = 0; k < CPT; ++k)
#pragma unroll
            for (int i = 0; i < NN; ++i)
              d[5] += (r[k][i] - y[k][i]) * (r[k][i] - y[k][i]);
          rr_warm = block_sum<W, 0>(d[5], red);
          alpha0 = 0.0;
          beta0 = 1.0;
        }
      else
        {
#pragma unroll
          for (int k = 0; k < CPT; ++k)
            if (own[k])
              {
                double Ag[NN];
                sipg2d<N>(Pt, n0, n1, cx[k], cy[k], tab, Ag);
#pragma unroll
                for (int i = 0; i < NN; ++i)
                  {
                    const double x = phit[(long long)i * nc + cl[k]];
                    const double g =
                      guess[(long long)cl[k] * NN + i] - mean_g;
                    d[0] += x * y[k][i];
                    d[1] += x * Ag[i];
                    d[2] += g * Ag[i];
                    d[3] += x * r[k][i];
                    d[4] += g * r[k][i];
                    d[5] += (r[k][i] - y[k][i]) * (r[k][i] - y[k][i]);
                  }
              }
          d[0] = block_sum<W, 1>(d[0], red);
          d[1] = block_sum<W, 0>(d[1], red);
          d[2] = block_sum<W, 1>(d[2], red);
          d[3] = block_sum<W, 0>(d[3], red);
          d[4] = block_sum<W, 1>(d[4], red);
          d[5] = block_sum<W, 0>(d[5], red);
          rr_warm = d[5];
          const double det = d[0] * d[2] - d[1] * d[1];
          if (d[0] > 0.0 && d[2] > 0.0 && det > 1e-12 * d[0] * d[2])
            {
              alpha0 = (d[3] * d[2] - d[4] * d[1]) / det;
              beta0 = (d[0] * d[4] - d[1] * d[3]) / det;
            }
          else if (!(d[0] > 0.0) && d[2] > 0.0)
            {
              alpha0 = 0.0;
              beta0 = d[4] / d[2];
            }
          if (!isfinite(alpha0) || !isfinite(beta0))
            {
              alpha0 = 1.0;
              beta0 = 0.0;
            }
        }
      // x = alpha phi + beta g, A x = alpha A phi + beta A g (A g again:
      // the direction buffer still holds g)
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
          {
            double Ag[NN];
            sipg2d<N>(Pt, n0, n1, cx[k], cy[k], tab, Ag);
#pragma unroll
            for (int i = 0; i < NN; ++i)
              {
                const long long gi = (long long)i * nc + cl[k];
                phit[gi] =
                  alpha0 * phit[gi] +
                  beta0 * (guess[(long long)cl[k] * NN + i] - mean_g);
                y[k][i] = alpha0 * y[k][i] + beta0 * Ag[i];
              }
          }
    }
  double rr = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k)
#pragma unroll
    for (int i = 0; i < NN; ++i)
      {
        r[k][i] -= y[k][i];
        rr += r[k][i] * r[k][i];
      }
  if (res_image)
    {
      // the residual of the start only (no solve), r = b - A phi, as the
      // image (r - mean r) / rms(r - mean r) of the x nodes (row = x1 node,
      // x0 node fastest) and its scale rms
      double sr = 0.0;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
#pragma unroll
          for (int i = 0; i < NN; ++i)
            sr += r[k][i];
      const double mean = block_sum<W, 1>(sr, red) / nx;
      double sq = 0.0;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
#pragma unroll
          for (int i = 0; i < NN; ++i)
            sq += (r[k][i] - mean) * (r[k][i] - mean);
      const double rms = sqrt(block_sum<W, 0>(sq, red) / nx);
      const double inv = rms > 0.0 ? 1.0 / rms : 0.0;
      const int row = n0 * N;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
#pragma unroll
          for (int iy = 0; iy < N; ++iy)
#pragma unroll
            for (int ix = 0; ix < N; ++ix)
              res_image[(cy[k] * N + iy) * row + cx[k] * N + ix] =
                (float)((r[k][ix + N * iy] - mean) * inv);
      if (t == 0)
        *res_scale = rms;
      return;
    }
  rr = block_sum<W, 1>(rr, red); // also: every thread is done with Pt
  const double rr_start = rr;
  // the iterate of this thread's cells stays in registers during the
  // iterations (no global read-modify-write on the critical path)
  double xk[CPT][NN];
#pragma unroll
  for (int k = 0; k < CPT; ++k)
#pragma unroll
    for (int i = 0; i < NN; ++i)
      xk[k][i] = own[k] ? phit[(long long)i * nc + cl[k]] : 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k)
    if (own[k])
#pragma unroll
      for (int i = 0; i < NN; ++i)
        Pt[i * kPStride + cl[k]] = r[k][i];
  const double stop = rel_tol * sqrt(bb);
  int it = 0;
  while (sqrt(rr) > stop && it < max_it)
    {
      __syncthreads(); // the new direction is visible
      double pAp = 0.0;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
          {
            sipg2d<N>(Pt, n0, n1, cx[k], cy[k], tab, y[k]);
#pragma unroll
            for (int i = 0; i < NN; ++i)
              pAp += Pt[i * kPStride + cl[k]] * y[k][i];
          }
      pAp = block_sum<W, 0>(pAp, red);
      if (!(pAp > 0.0))
        break; // what is left of the residual lies in the null space
      const double alpha = rr / pAp;
      double rr2 = 0.0;
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
#pragma unroll
          for (int i = 0; i < NN; ++i)
            {
              xk[k][i] += alpha * Pt[i * kPStride + cl[k]];
              r[k][i] -= alpha * y[k][i];
              rr2 += r[k][i] * r[k][i];
            }
      rr2 = block_sum<W, 1>(rr2, red);
      const double beta = rr2 / rr;
      rr = rr2;
      // every matrix-vector product of this iteration ended before the
      // barrier in the last sum: the direction can be overwritten
#pragma unroll
      for (int k = 0; k < CPT; ++k)
        if (own[k])
#pragma unroll
          for (int i = 0; i < NN; ++i)
            Pt[i * kPStride + cl[k]] =
              r[k][i] + beta * Pt[i * kPStride + cl[k]];
      ++it;
    }

  // phi back in the cell-major layout, then grad phi at the quadrature
  // points of this thread's cells
#pragma unroll
  for (int k = 0; k < CPT; ++k)
    if (own[k])
      {
        double x[NN], g[NN];
#pragma unroll
        for (int i = 0; i < NN; ++i)
          x[i] = xk[k][i];
#pragma unroll
        for (int i = 0; i < NN; ++i)
          phi[(long long)cl[k] * NN + i] = x[i];
        // d/dx0: G along x0, S along x1; d/dx1: S along x0, G along x1
#pragma unroll
        for (int e = 0; e < 2; ++e)
          {
#pragma unroll
            for (int iy = 0; iy < N; ++iy)
#pragma unroll
              for (int qx = 0; qx < N; ++qx)
                {
                  double a = 0.0;
#pragma unroll
                  for (int jx = 0; jx < N; ++jx)
                    a += (e == 0 ? xd.G[qx * N + jx] : xd.S[qx * N + jx]) *
                         x[jx + N * iy];
                  g[qx + N * iy] = a;
                }
#pragma unroll
            for (int qy = 0; qy < N; ++qy)
#pragma unroll
              for (int qx = 0; qx < N; ++qx)
                {
                  double a = 0.0;
#pragma unroll
                  for (int jy = 0; jy < N; ++jy)
                    a += (e == 1 ? xd.G[qy * N + jy] : xd.S[qy * N + jy]) *
                         g[qx + N * jy];
                  efield[((long long)cl[k] * NN + qx + N * qy) * 2 + e] =
                    a / xd.h[e];
                }
          }
      }
  if (threadIdx.x == 0)
    {
      double *ds = reinterpret_cast<double *>(stats);
      const double nb = sqrt(bb) + 1e-300;
      stats[0] += it;
      stats[1] += 1;
      ds[2] = sqrt(rr) / nb;
      ds[3] = sqrt(rr_start) / nb;
      ds[4] = (guess ? sqrt(rr_warm) : sqrt(rr_start)) / nb;
      if (guess)
        {
          stats[5] += 1;
          ds[6] = alpha0;
          ds[7] = beta0;
        }
    }
}

// ---- the 2+2 phase-space apply: hd_vpwave.hip ----

// ---- diagnostics: mass, ||f||^2, kinetic energy, momentum (3), per cell ----
__global__ __launch_bounds__(256) void moments_kernel(
  const double *__restrict__ f, MeshDev mesh, NormDev nt, int dim_x,
  double *__restrict__ partial)
{
  extern __shared__ double smem[];
  const int D = mesh.dim, nd = nt.nd, nq = nt.nq;
  int dofs = 1, vol = 1;
  for (int d = 0; d < D; ++d)
    {
      dofs *= nd;
      vol *= nq;
    }
  double *b0 = smem, *b1 = smem + vol, *red = smem + 2 * vol;
  const long long cell = blockIdx.x;
  for (int i = threadIdx.x; i < dofs; i += blockDim.x)
    b0[i] = f[cell * dofs + i];
  __syncthreads();
  double *in = b0, *out = b1;
  for (int d = 0; d < D; ++d)
    {
      int pre = 1, post = 1;
      for (int j = 0; j < d; ++j)
        pre *= nq;
      for (int j = d + 1; j < D; ++j)
        post *= nd;
      for (int o = threadIdx.x; o < pre * nq * post; o += blockDim.x)
        {
          const int pi = o % pre, rest = o / pre;
          const int q = rest % nq, po = rest / nq;
          double acc = 0.0;
          for (int j = 0; j < nd; ++j)
            acc += nt.S[q * nd + j] * in[pi + pre * (j + nd * po)];
          out[o] = acc;
        }
      __syncthreads();
      double *t = in;
      in = out;
      out = t;
    }
  double m[6] = {0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < vol; q += blockDim.x)
    {
      double jxw = 1.0, v[3] = {0, 0, 0}, v2 = 0.0;
      long long c = cell;
      int r = q;
      for (int d = 0; d < D; ++d)
        {
          const int cd = (int)(c % mesh.n[d]);
          c /= mesh.n[d];
          const int qd = r % nq;
          r /= nq;
          jxw *= mesh.h[d] * nt.w[qd];
          if (d >= dim_x)
            {
              v[d - dim_x] = mesh.lo[d] + (cd + nt.qp[qd]) * mesh.h[d];
              v2 += v[d - dim_x] * v[d - dim_x];
            }
        }
      const double fv = in[q];
      m[0] += fv * jxw;
      m[1] += fv * fv * jxw;
      m[2] += v2 * fv * jxw;
      for (int e = 0; e < 3; ++e)
        m[3 + e] += v[e] * fv * jxw;
    }
  for (int k = 0; k < 6; ++k)
    {
      red[threadIdx.x] = m[k];
      __syncthreads();
      for (int s = blockDim.x / 2; s > 0; s /= 2)
        {
          if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
          __syncthreads();
        }
      if (threadIdx.x == 0)
        partial[cell * 6 + k] = red[0];
      __syncthreads();
    }
}
} // namespace vlasov

hipError_t vp_density(const double *f, const float *f32, const VpDev &vp,
                      long long n_x_cells, long long n_v_cells,
                      double *partial, int n_chunks, double *rho,
                      hipStream_t stream, float *copy32, double *copy64)
{
  const int chunk = (int)((n_v_cells + n_chunks - 1) / n_chunks);
  // a multiple of the x DoFs per cell, so each thread keeps one x node
  const int threads = (256 / vp.dofs_x) * vp.dofs_x;
  hipLaunchKernelGGL(vlasov::density_partial,
                     dim3((unsigned)n_x_cells, (unsigned)n_chunks),
                     dim3(threads),
                     0, stream, f, f32, vp, n_x_cells, n_v_cells, chunk,
                     partial, copy32, copy64);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  const long long nx = n_x_cells * vp.dofs_x;
  hipLaunchKernelGGL(vlasov::density_reduce, dim3((unsigned)((nx + 255) / 256)),
                     dim3(256), 0, stream, partial, n_chunks, vp.dofs_x, nx,
                     rho);
  return hipGetLastError();
}

bool vp_poisson_supported(int dim_x, int N, long long n_x_cells)
{
  const long long dofs = n_x_cells * ipow(N, dim_x);
  if ((dofs + vlasov::kSolverThreads) * 8 > 150 * 1024)
    return false;
  return (dim_x == 1 && N <= 8) || (dim_x == 2 && N <= 6) ||
         (dim_x == 3 && N <= 4);
}

bool vp_poisson2d_supported(int dim_x, int N, long long n_x_cells)
{
  // the direction in LDS: N^2 rows of kPStride doubles (128 KB at N = 4)
  return dim_x == 2 && N >= 2 && N <= 4 &&
         n_x_cells <= vlasov::kSolverThreads;
}

void vp_poisson2d_table(const XDev &xd, int N, double *tab)
{
  // L_d (left, self, right) and M_d per direction, from the 1D mass M,
  // stiffness K and end-point derivatives d0, d1 of the reference cell;
  // SIPG penalty sigma = k (k + 1) * 2 / h (poisson.h:193-254)
  const int NN = N * N;
  for (int d = 0; d < 2; ++d)
    {
      const double h = xd.h[d], sg = xd.kk * (2.0 / h);
      double *Ll = tab + (3 * d) * NN, *Ls = Ll + NN, *Lr = Ls + NN;
      double *M = tab + (6 + d) * NN;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          {
            const double di0 = xd.d0[i] / h, di1 = xd.d1[i] / h;
            const double dj0 = xd.d0[j] / h, dj1 = xd.d1[j] / h;
            const bool i0 = i == 0, iN = i == N - 1, j0 = j == 0,
                       jN = j == N - 1;
            // cell stiffness; the right face (this cell minus), the left
            // face (this cell plus): sigma [u][v] - {u'}[v] - [u]{v'}
            Ls[i * N + j] = xd.K[i * N + j] / h +
                            (iN ? sg * jN - 0.5 * dj1 : 0.0) -
                            (jN ? 0.5 * di1 : 0.0) +
                            (i0 ? sg * j0 + 0.5 * dj0 : 0.0) +
                            (j0 ? 0.5 * di0 : 0.0);
            Lr[i * N + j] = (iN ? -sg * j0 - 0.5 * dj0 : 0.0) +
                            (j0 ? 0.5 * di1 : 0.0);
            Ll[i * N + j] = (i0 ? -(sg * jN - 0.5 * dj1) : 0.0) -
                            (jN ? 0.5 * di0 : 0.0);
            M[i * N + j] = xd.M[i * N + j] * h;
          }
    }
}

hipError_t vp_poisson(const XDev &xd, int N, const double *rho, double *phi,
                      const double *guess, double *efield, double *work,
                      double rel_tol, int max_it, long long *stats,
                      hipStream_t stream, const double *tab2d,
                      const ResidualImage *residual, bool guess_alone)
{
  if (residual &&
      !(tab2d && vp_poisson2d_supported(xd.dim, N, xd.n_cells)))
    return hipErrorInvalidValue;
  if (tab2d && vp_poisson2d_supported(xd.dim, N, xd.n_cells))
    {
      const size_t lds =
        (size_t)(vlasov::kPStride * N * N + 32) * sizeof(double);
#define HD_VP2(NN_)                                                            \
  if (N == NN_)                                                                \
    {                                                                          \
      constexpr int CPT = 2;                                                   \
      hipError_t e = hipFuncSetAttribute(                                      \
        reinterpret_cast<const void *>(                                        \
          &vlasov::poisson2d_kernel<NN_, CPT>),                                \
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                 \
      if (e != hipSuccess)                                                     \
        return e;                                                              \
      hipLaunchKernelGGL((vlasov::poisson2d_kernel<NN_, CPT>), dim3(1),        \
                         dim3(vlasov::kSolverThreads / CPT), lds, stream, xd,  \
                         tab2d, rho, phi, guess, efield, work, rel_tol,        \
                         max_it, stats,                                        \
                         residual ? residual->image : nullptr,                 \
                         residual ? residual->scale : nullptr,                 \
                         guess_alone ? 1 : 0);                                 \
      return hipGetLastError();                                                \
    }
      HD_VP2(2) HD_VP2(3) HD_VP2(4)
#undef HD_VP2
    }
  const long long dofs = xd.n_cells * ipow(N, xd.dim);
  const size_t lds = (size_t)(dofs + vlasov::kSolverThreads) * sizeof(double);
#define HD_VP(DX, NN_)                                                         \
  if (xd.dim == DX && N == NN_)                                                \
    {                                                                          \
      hipError_t e = hipFuncSetAttribute(                                      \
        reinterpret_cast<const void *>(&vlasov::poisson_kernel<DX, NN_>),      \
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                 \
      if (e != hipSuccess)                                                     \
        return e;                                                              \
      hipLaunchKernelGGL((vlasov::poisson_kernel<DX, NN_>), dim3(1),           \
                         dim3(vlasov::kSolverThreads), lds, stream, xd, rho,   \
                         phi, guess, efield, work, rel_tol, max_it, stats);    \
      return hipGetLastError();                                                \
    }
  HD_VP(1, 2) HD_VP(1, 3) HD_VP(1, 4) HD_VP(1, 5) HD_VP(1, 6) HD_VP(1, 7)
  HD_VP(1, 8) HD_VP(2, 2) HD_VP(2, 3) HD_VP(2, 4) HD_VP(2, 5) HD_VP(2, 6)
  HD_VP(3, 2) HD_VP(3, 3) HD_VP(3, 4)
#undef HD_VP
  return hipErrorInvalidValue;
}

hipError_t vp_moments(const double *f, const MeshDev &mesh, const NormDev &nt,
                      int dim_x, int vol, double *partial, hipStream_t stream)
{
  const size_t lds = (2 * (size_t)vol + 256) * sizeof(double);
  hipError_t e = hipFuncSetAttribute(
    reinterpret_cast<const void *>(&vlasov::moments_kernel),
    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess)
    return e;
  hipLaunchKernelGGL(vlasov::moments_kernel, dim3((unsigned)mesh.n_cells),
                     dim3(256), lds, stream, f, mesh, nt, dim_x, partial);
  return hipGetLastError();
}

namespace vlasov
{
// rho[i] = sum over v cells of part[vc][i].  A block takes kDensX x nodes;
// its thread (x, g) sums the v cells g, g + kDensG, ... in order, then the
// kDensG partial sums are added in group order (a fixed order: the result
// does not depend on the launch).  Many blocks and independent loads per
// thread keep enough reads in flight (one thread per x node, walking all v
// cells, left the GPU idle).
constexpr int kDensX = 32, kDensG = 8;
__global__ __launch_bounds__(kDensX *kDensG) void density_from_partials(
  const double *__restrict__ part, long long n_v_cells, long long n_x_dofs,
  double *__restrict__ rho)
{
  __shared__ double sum[kDensG][kDensX];
  const int x = threadIdx.x % kDensX, g = threadIdx.x / kDensX;
  const long long i = (long long)blockIdx.x * kDensX + x;
  double s = 0.0;
  if (i < n_x_dofs)
    {
#pragma unroll 8
      for (long long vc = g; vc < n_v_cells; vc += kDensG)
        s += part[vc * n_x_dofs + i];
    }
  sum[g][x] = s;
  __syncthreads();
  if (g == 0 && i < n_x_dofs)
    {
      double r = sum[0][x];
#pragma unroll
      for (int k = 1; k < kDensG; ++k)
        r += sum[k][x];
      rho[i] = r;
    }
}
} // namespace vlasov

hipError_t vp_density_from_partials(const double *part, long long n_v_cells,
                                    long long n_x_dofs, double *rho,
                                    hipStream_t stream)
{
  hipLaunchKernelGGL(
    vlasov::density_from_partials,
    dim3((unsigned)((n_x_dofs + vlasov::kDensX - 1) / vlasov::kDensX)),
    dim3(vlasov::kDensX * vlasov::kDensG), 0, stream, part, n_v_cells,
    n_x_dofs, rho);
  return hipGetLastError();
}

bool vp_apply_supported(int dim_x, int dim_v, int N)
{
  return dim_x == 2 && dim_v == 2 && N >= 2 && N <= 5;
}

int vp_apply_table_size(int N)
{
  switch (N)
    {
      case 2:
        return vlasov::vptab::Tab<2>::size;
      case 3:
        return vlasov::vptab::Tab<3>::size;
      case 4:
        return vlasov::vptab::Tab<4>::size;
      case 5:
        return vlasov::vptab::Tab<5>::size;
      default:
        return 0;
    }
}

} // namespace hd
