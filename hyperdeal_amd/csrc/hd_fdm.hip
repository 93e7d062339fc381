// hd_fdm.hip -- an exact start for the 2D periodic SIPG Poisson solve by
// fast diagonalisation.
//
// On a periodic Cartesian x mesh the SIPG Laplacian of the Vlasov-Poisson
// field solve (hyper.deal: examples/vlasov_poisson/include/poisson.h:
// 100-312) is separable: with X[g1][g0] the nodal values (g_d = c_d N + i_d)
//   A X = M1 X L0^T + L1 X M0^T,
// L_d the 1D periodic SIPG operator (blocks left / self / right of the 2D
// solver's table) and M_d the 1D mass matrix (block diagonal).  With the
// generalised eigenvectors L_d V_d = M_d V_d Lambda_d, V_d^T M_d V_d = I,
//   X = V1 [ (V1^T B V0) / (lambda1_i + lambda0_j) ] V0^T
// solves A X = B exactly, the constant mode (lambda = 0) left out.  The
// eigenvectors come from a Jacobi eigensolver on the host at set-up; a solve
// is four dense products of (n_d N) x (n_d N) matrices on the GPU.  The CG
// solver then starts from this X and, as for any start, runs until the
// relative residual is below its tolerance (with an exact start: none or
// one iteration).
#include "hd_device.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace hd
{
namespace fdm
{
constexpr int kTile = 16;
constexpr int kRhsThreads = 1024;

// Rhs as the CG solver forms it: B = -(M0 (x) M1)(rho - mean rho), then
// B -= mean B; written as B[g1][g0].  One block.
template <int N>
__global__ __launch_bounds__(kRhsThreads) void rhs_kernel(
  const double *__restrict__ tab, const double *__restrict__ rho, int n0,
  int n1, double *__restrict__ B)
{
  constexpr int NN = N * N;
  __shared__ double red[kRhsThreads / 64];
  const int nc = n0 * n1, P0 = n0 * N;
  const double nx = (double)nc * NN;
  auto block_sum = [&](double v) {
    for (int off = 32; off >= 1; off >>= 1)
      v += __shfl_xor(v, off, 64);
    __syncthreads(); // earlier readers of red are done
    if ((threadIdx.x & 63) == 0)
      red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < kRhsThreads / 64; ++w)
      r += red[w];
    return r;
  };
  double s = 0.0;
  for (int i = threadIdx.x; i < nc * NN; i += kRhsThreads)
    s += rho[i];
  const double mean_rho = block_sum(s) / nx;
  s = 0.0;
  for (int c = threadIdx.x; c < nc; c += kRhsThreads)
    {
      const int cx = c % n0, cy = c / n0;
      double y[NN];
      for (int iy = 0; iy < N; ++iy)
        for (int ix = 0; ix < N; ++ix)
          {
            double a = 0.0;
            for (int jx = 0; jx < N; ++jx)
              a += tab[6 * NN + ix * N + jx] *
                   (rho[(long long)c * NN + jx + N * iy] - mean_rho);
            y[ix + N * iy] = a;
          }
      for (int iy = 0; iy < N; ++iy)
        for (int ix = 0; ix < N; ++ix)
          {
            double a = 0.0;
            for (int jy = 0; jy < N; ++jy)
              a += tab[7 * NN + iy * N + jy] * y[ix + N * jy];
            B[(long long)(cy * N + iy) * P0 + cx * N + ix] = -a;
            s -= a;
          }
    }
  const double mean_b = block_sum(s) / nx;
  // each thread subtracts from the values it wrote
  for (int c = threadIdx.x; c < nc; c += kRhsThreads)
    {
      const int cx = c % n0, cy = c / n0;
      for (int iy = 0; iy < N; ++iy)
        for (int ix = 0; ix < N; ++ix)
          B[(long long)(cy * N + iy) * P0 + cx * N + ix] -= mean_b;
    }
}

// One product of the solve, one thread per output (r, c), r < P1, c < P0:
//   STEP 0: T = B V0            (k < P0)
//   STEP 1: Y = inv .* (V1^T T) (k < P1)
//   STEP 2: Z = V1 Y            (k < P1)
//   STEP 3: X = Z V0^T          (k < P0), X written cell-major
// The k-loop runs over 16-wide tiles of both operands staged in LDS.
template <int STEP, int N>
__global__ __launch_bounds__(kTile *kTile) void product_kernel(
  const double *__restrict__ in, const double *__restrict__ V0,
  const double *__restrict__ V1, const double *__restrict__ inv, int n0,
  int n1, double *__restrict__ out)
{
  __shared__ double a[kTile][kTile + 1], b[kTile][kTile + 1];
  const int P0 = n0 * N, P1 = n1 * N;
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int r = blockIdx.y * kTile + ty, c = blockIdx.x * kTile + tx;
  const int K = (STEP == 0 || STEP == 3) ? P0 : P1;
  double acc = 0.0;
  for (int k0 = 0; k0 < K; k0 += kTile)
    {
      // a[ty][kk] = left operand (r, k0 + kk); b[kk][tx] = right (k0 + kk, c)
      {
        const int ra = blockIdx.y * kTile + ty, ka = k0 + tx;
        double va = 0.0;
        if (ra < P1 && ka < K)
          {
            if constexpr (STEP == 0)
              va = in[(long long)ra * P0 + ka]; // B[r][k]
            else if constexpr (STEP == 1)
              va = V1[(long long)ka * P1 + ra]; // V1^T[r][k]
            else if constexpr (STEP == 2)
              va = V1[(long long)ra * P1 + ka]; // V1[r][k]
            else
              va = in[(long long)ra * P0 + ka]; // Z[r][k]
          }
        a[ty][tx] = va;
        const int kb = k0 + ty, cb = blockIdx.x * kTile + tx;
        double vb = 0.0;
        if (kb < K && cb < P0)
          {
            if constexpr (STEP == 0)
              vb = V0[(long long)kb * P0 + cb]; // V0[k][c]
            else if constexpr (STEP == 1 || STEP == 2)
              vb = in[(long long)kb * P0 + cb]; // T[k][c], Y[k][c]
            else
              vb = V0[(long long)cb * P0 + kb]; // V0^T[k][c]
          }
        b[ty][tx] = vb;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < kTile; ++kk)
        acc += a[ty][kk] * b[kk][tx];
      __syncthreads();
    }
  if (r >= P1 || c >= P0)
    return;
  if constexpr (STEP == 1)
    acc *= inv[(long long)r * P0 + c];
  if constexpr (STEP == 3)
    {
      const int cx = c / N, ix = c % N, cy = r / N, iy = r % N;
      out[((long long)(cy * n0 + cx)) * (N * N) + ix + N * iy] = acc;
    }
  else
    out[(long long)r * P0 + c] = acc;
}

// Symmetric eigenproblem C = Q diag(lam) Q^T by cyclic Jacobi rotations
// (C overwritten), n x n row-major.
void jacobi_eigen(std::vector<double> &C, int n, std::vector<double> &lam,
                  std::vector<double> &Q)
{
  Q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    Q[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep)
    {
      double off = 0.0, diag = 0.0;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          (i == j ? diag : off) += C[(size_t)i * n + j] * C[(size_t)i * n + j];
      if (off <= 1e-30 * diag)
        break;
      for (int p = 0; p < n - 1; ++p)
        for (int q = p + 1; q < n; ++q)
          {
            const double apq = C[(size_t)p * n + q];
            if (std::fabs(apq) < 1e-300)
              continue;
            const double app = C[(size_t)p * n + p], aqq = C[(size_t)q * n + q];
            // rotation angle: tan(2 theta) = 2 apq / (aqq - app)
            const double theta = 0.5 * (aqq - app) / apq;
            const double t = (theta >= 0 ? 1.0 : -1.0) /
                             (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
            for (int k = 0; k < n; ++k)
              {
                const double ckp = C[(size_t)k * n + p], ckq = C[(size_t)k * n + q];
                C[(size_t)k * n + p] = cs * ckp - sn * ckq;
                C[(size_t)k * n + q] = sn * ckp + cs * ckq;
              }
            for (int k = 0; k < n; ++k)
              {
                const double cpk = C[(size_t)p * n + k], cqk = C[(size_t)q * n + k];
                C[(size_t)p * n + k] = cs * cpk - sn * cqk;
                C[(size_t)q * n + k] = sn * cpk + cs * cqk;
              }
            for (int k = 0; k < n; ++k)
              {
                const double qkp = Q[(size_t)k * n + p], qkq = Q[(size_t)k * n + q];
                Q[(size_t)k * n + p] = cs * qkp - sn * qkq;
                Q[(size_t)k * n + q] = sn * qkp + cs * qkq;
              }
          }
    }
  lam.resize(n);
  for (int i = 0; i < n; ++i)
    lam[i] = C[(size_t)i * n + i];
}

// V (P x P, P = n N) with L V = M V diag(lam), V^T M V = I for the 1D
// periodic operator of direction d; false if M is not positive definite
bool direction_basis(const double *tab, int N, int d, int n,
                     std::vector<double> &V, std::vector<double> &lam)
{
  const int NN = N * N, P = n * N;
  const double *Ll = tab + 3 * d * NN, *Ls = Ll + NN, *Lr = Ls + NN;
  const double *Mb = tab + (6 + d) * NN;
  std::vector<double> L((size_t)P * P, 0.0);
  for (int c = 0; c < n; ++c)
    {
      const int cl = (c + n - 1) % n, cr = (c + 1) % n;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          {
            L[(size_t)(c * N + i) * P + cl * N + j] += Ll[i * N + j];
            L[(size_t)(c * N + i) * P + c * N + j] += Ls[i * N + j];
            L[(size_t)(c * N + i) * P + cr * N + j] += Lr[i * N + j];
          }
    }
  // Cholesky of the N x N mass block: Mb = R^T R, R upper triangular
  std::vector<double> R((size_t)NN, 0.0), Ri((size_t)NN, 0.0);
  for (int j = 0; j < N; ++j)
    {
      double s = Mb[j * N + j];
      for (int k = 0; k < j; ++k)
        s -= R[k * N + j] * R[k * N + j];
      if (!(s > 0.0))
        return false;
      R[j * N + j] = std::sqrt(s);
      for (int i = j + 1; i < N; ++i)
        {
          double t = Mb[j * N + i];
          for (int k = 0; k < j; ++k)
            t -= R[k * N + j] * R[k * N + i];
          R[j * N + i] = t / R[j * N + j];
        }
    }
  // Ri = R^{-1} (upper triangular)
  for (int j = N - 1; j >= 0; --j)
    {
      Ri[j * N + j] = 1.0 / R[j * N + j];
      for (int i = j - 1; i >= 0; --i)
        {
          double s = 0.0;
          for (int k = i + 1; k <= j; ++k)
            s += R[i * N + k] * Ri[k * N + j];
          Ri[i * N + j] = -s / R[i * N + i];
        }
    }
  // C = Rb^-T L Rb^-1 with Rb = blockdiag(R): per block pair (a, b)
  // C_ab = Ri^T L_ab Ri
  std::vector<double> C((size_t)P * P, 0.0), tmp((size_t)P * P, 0.0);
  for (int r = 0; r < P; ++r) // tmp = L Rb^-1
    for (int cb = 0; cb < n; ++cb)
      for (int j = 0; j < N; ++j)
        {
          double s = 0.0;
          for (int k = 0; k <= j; ++k)
            s += L[(size_t)r * P + cb * N + k] * Ri[k * N + j];
          tmp[(size_t)r * P + cb * N + j] = s;
        }
  for (int ra = 0; ra < n; ++ra) // C = Rb^-T tmp
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < P; ++c)
        {
          double s = 0.0;
          for (int k = 0; k <= i; ++k)
            s += Ri[k * N + i] * tmp[(size_t)(ra * N + k) * P + c];
          C[(size_t)(ra * N + i) * P + c] = s;
        }
  for (int i = 0; i < P; ++i) // symmetric up to roundoff
    for (int j = i + 1; j < P; ++j)
      {
        const double m = 0.5 * (C[(size_t)i * P + j] + C[(size_t)j * P + i]);
        C[(size_t)i * P + j] = C[(size_t)j * P + i] = m;
      }
  std::vector<double> Q;
  jacobi_eigen(C, P, lam, Q);
  // V = Rb^-1 Q
  V.assign((size_t)P * P, 0.0);
  for (int ra = 0; ra < n; ++ra)
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < P; ++c)
        {
          double s = 0.0;
          for (int k = i; k < N; ++k)
            s += Ri[i * N + k] * Q[(size_t)(ra * N + k) * P + c];
          V[(size_t)(ra * N + i) * P + c] = s;
        }
  return true;
}
} // namespace fdm

bool vp_fdm_supported(int dim_x, int N, const int *n, long long n_x_cells)
{
  return vp_poisson2d_supported(dim_x, N, n_x_cells) && n[0] * N <= 256 &&
         n[1] * N <= 256;
}

hipError_t vp_fdm_create(const double *tab_host, int N, int n0, int n1,
                         FdmDev *out)
{
  *out = FdmDev{};
  std::vector<double> V0, V1, l0, l1;
  if (!fdm::direction_basis(tab_host, N, 0, n0, V0, l0) ||
      !fdm::direction_basis(tab_host, N, 1, n1, V1, l1))
    return hipErrorInvalidValue;
  const int P0 = n0 * N, P1 = n1 * N;
  double lmax = 0.0;
  for (double v : l0)
    lmax = std::max(lmax, std::fabs(v));
  for (double v : l1)
    lmax = std::max(lmax, std::fabs(v));
  std::vector<double> inv((size_t)P1 * P0);
  for (int r = 0; r < P1; ++r)
    for (int c = 0; c < P0; ++c)
      {
        const double s = l1[r] + l0[c];
        // the constant mode (lambda0 = lambda1 = 0) is left out
        inv[(size_t)r * P0 + c] = s > 1e-9 * lmax ? 1.0 / s : 0.0;
      }
  const size_t n_all = (size_t)P0 * P0 + (size_t)P1 * P1 +
                       3 * (size_t)P1 * P0 + (size_t)n0 * n1 * N * N;
  double *mem = nullptr;
  hipError_t e = hipMalloc(&mem, n_all * sizeof(double));
  if (e != hipSuccess)
    return e;
  FdmDev f{};
  f.mem = mem;
  f.V0 = mem;
  f.V1 = f.V0 + (size_t)P0 * P0;
  f.inv = f.V1 + (size_t)P1 * P1;
  f.B = f.inv + (size_t)P1 * P0;
  f.T = f.B + (size_t)P1 * P0;
  f.x = f.T + (size_t)P1 * P0;
  f.N = N;
  f.n0 = n0;
  f.n1 = n1;
  e = hipMemcpy(f.V0, V0.data(), V0.size() * sizeof(double),
                hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(f.V1, V1.data(), V1.size() * sizeof(double),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(f.inv, inv.data(), inv.size() * sizeof(double),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess)
    {
      (void)hipFree(mem);
      return e;
    }
  *out = f;
  return hipSuccess;
}

void vp_fdm_destroy(FdmDev &f)
{
  (void)hipFree(f.mem);
  f = FdmDev{};
}

hipError_t vp_fdm_solve(const FdmDev &f, const double *tab, const double *rho,
                        hipStream_t stream)
{
  const int N = f.N, n0 = f.n0, n1 = f.n1, P0 = n0 * N, P1 = n1 * N;
  const dim3 grid((P0 + fdm::kTile - 1) / fdm::kTile,
                  (P1 + fdm::kTile - 1) / fdm::kTile);
  const dim3 block(fdm::kTile * fdm::kTile);
#define HD_FDM(NN_)                                                          \
  if (N == NN_)                                                              \
    {                                                                        \
      hipLaunchKernelGGL(fdm::rhs_kernel<NN_>, dim3(1),                      \
                         dim3(fdm::kRhsThreads), 0, stream, tab, rho, n0, n1, \
                         f.B);                                               \
      hipLaunchKernelGGL((fdm::product_kernel<0, NN_>), grid, block, 0,      \
                         stream, f.B, f.V0, f.V1, f.inv, n0, n1, f.T);       \
      hipLaunchKernelGGL((fdm::product_kernel<1, NN_>), grid, block, 0,      \
                         stream, f.T, f.V0, f.V1, f.inv, n0, n1, f.B);       \
      hipLaunchKernelGGL((fdm::product_kernel<2, NN_>), grid, block, 0,      \
                         stream, f.B, f.V0, f.V1, f.inv, n0, n1, f.T);       \
      hipLaunchKernelGGL((fdm::product_kernel<3, NN_>), grid, block, 0,      \
                         stream, f.T, f.V0, f.V1, f.inv, n0, n1, f.x);       \
      return hipGetLastError();                                              \
    }
  HD_FDM(2) HD_FDM(3) HD_FDM(4)
#undef HD_FDM
  return hipErrorInvalidValue;
}
} // namespace hd
