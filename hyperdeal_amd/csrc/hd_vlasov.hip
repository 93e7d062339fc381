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

namespace hd
{
namespace vlasov
{
constexpr int kSolverThreads = 1024;

// ---- rho: partial sums over chunks of v cells, then a fixed-order sum ----
__global__ __launch_bounds__(256) void density_partial(
  const double *__restrict__ f, const float *__restrict__ f32,
  VpDev vp, long long n_x_cells, long long n_v_cells, int chunk,
  double *__restrict__ partial)
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
                      hipStream_t stream)
{
  const int chunk = (int)((n_v_cells + n_chunks - 1) / n_chunks);
  // a multiple of the x DoFs per cell, so each thread keeps one x node
  const int threads = (256 / vp.dofs_x) * vp.dofs_x;
  hipLaunchKernelGGL(vlasov::density_partial,
                     dim3((unsigned)n_x_cells, (unsigned)n_chunks),
                     dim3(threads),
                     0, stream, f, f32, vp, n_x_cells, n_v_cells, chunk,
                     partial);
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

hipError_t vp_poisson(const XDev &xd, int N, const double *rho, double *phi,
                      const double *guess, double *efield, double *work,
                      double rel_tol, int max_it, long long *stats,
                      hipStream_t stream)
{
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
} // namespace hd
