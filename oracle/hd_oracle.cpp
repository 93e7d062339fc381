// hd_oracle.cpp -- CPU oracle of the phase-space DG advection operator.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, to check results or as
// the timed CPU baseline; the product path never does.
//
// What is restated here, and from where (paths relative to the hyper.deal
// source tree):
//   ECL fused operator ........ include/hyper.deal/operators/advection/
//                               advection_operation.h:221-566
//   FCL cell/face/boundary .... advection_operation.h:573-1131
//   evaluator geometry ........ matrix_free/fe_evaluation_cell.h:196-348,
//                               fe_evaluation_face.h:381-463,
//                               fe_evaluation_cell_inverse.h:123-134
//   face trace / lift ......... matrix_free/evaluation_kernels.h:64-330
//   LSRK integrators .......... base/time_integrators.templates.h:22-184
//   time loop ................. source/base/time_loop.cc:33-67
//   CFL ....................... operators/advection/cfl.h:56-124
//   dt rule + diagnostics ..... examples/advection/include/application.h:
//                               363-396, 456-506
//   interpolate / L2 norm ..... numerics/vector_tools.h:32-223
// The tensor-product kernels that hyper.deal takes from deal.II (shape
// values S, collocation gradients D, inverse shape values) are rebuilt from
// the mathematics: Lagrange bases on Gauss-Lobatto nodes and Gauss points.
//
// Parity pins: the golden diagnostics of examples/advection/tests (every
// ECL/FCL, periodic/Dirichlet, skew, collocation, over-integration and RK
// variant), tests/time_discretization/time_integrators_02 and
// tests/vector_tools/vector_tools_01, plus closed-form table checks and
// operator invariants in tests/test_oracle_*.py.
//
// Layout: lane-batched like hyper.deal's VectorizedArray macro cells: L
// consecutive cells are processed together with the lane index innermost so
// the compiler vectorises every contraction; OpenMP over macro cells replaces
// MPI ranks.

#include "hd_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace
{
constexpr int L = 8; // lanes per macro cell
constexpr double PI = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// 1D polynomial machinery
// ---------------------------------------------------------------------------

// Legendre P_n and its derivative at z (three-term recurrence).
void legendre(int n, long double z, long double &p, long double &dp,
              long double &pm1)
{
  long double p0 = 1.0L, p1 = z;
  if (n == 0)
    {
      p = 1.0L;
      dp = 0.0L;
      pm1 = 0.0L;
      return;
    }
  for (int j = 2; j <= n; ++j)
    {
      long double p2 = ((2.0L * j - 1.0L) * z * p1 - (j - 1.0L) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
  p = p1;
  pm1 = p0;
  dp = n * (z * p1 - p0) / (z * z - 1.0L);
}

// Gauss-Legendre rule with n points on [0,1], ascending.
void gauss_rule(int n, std::vector<double> &x, std::vector<double> &w)
{
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i)
    {
      long double z = std::cos(PI * (i + 0.75) / (n + 0.5));
      long double p, dp, pm1;
      for (int it = 0; it < 100; ++it)
        {
          legendre(n, z, p, dp, pm1);
          long double dz = p / dp;
          z -= dz;
          if (std::fabs((double)dz) < 1e-19)
            break;
        }
      legendre(n, z, p, dp, pm1);
      // ascending order on [0,1]: z_i is descending in i
      x[n - 1 - i] = (double)(0.5L * (1.0L + z));
      w[n - 1 - i] = (double)(1.0L / ((1.0L - z * z) * dp * dp));
    }
}

// Gauss-Lobatto rule with n >= 2 points on [0,1], ascending.
void gauss_lobatto_rule(int n, std::vector<double> &x, std::vector<double> &w)
{
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const int m = n - 1; // interior points are the roots of P'_m
  std::vector<long double> z(n);
  z[0] = -1.0L;
  z[n - 1] = 1.0L;
  for (int i = 1; i < n - 1; ++i)
    {
      long double zi = -std::cos(PI * i / m);
      for (int it = 0; it < 100; ++it)
        {
          long double p, dp, pm1;
          legendre(m, zi, p, dp, pm1);
          // P''_m from the Legendre ODE: (1-z^2)P'' = 2 z P' - m(m+1) P
          long double ddp = (2.0L * zi * dp - m * (m + 1.0L) * p) /
                            (1.0L - zi * zi);
          long double dz = dp / ddp;
          zi -= dz;
          if (std::fabs((double)dz) < 1e-19)
            break;
        }
      z[i] = zi;
    }
  for (int i = 0; i < n; ++i)
    {
      long double p, dp, pm1;
      legendre(m, z[i], p, dp, pm1);
      if (i == 0)
        p = (m % 2 == 0) ? 1.0L : -1.0L;
      if (i == n - 1)
        p = 1.0L;
      x[i] = (double)(0.5L * (1.0L + z[i]));
      w[i] = (double)(1.0L / (n * (n - 1.0L) * p * p)); // 2/(n(n-1)P^2) / 2
    }
}

// Lagrange basis on nodes xs: value of l_j at x.
long double lagrange_value(const std::vector<double> &xs, int j, long double x)
{
  long double r = 1.0L;
  for (size_t m = 0; m < xs.size(); ++m)
    if ((int)m != j)
      r *= (x - xs[m]) / ((long double)xs[j] - xs[m]);
  return r;
}

// Derivative of l_j at x.
long double lagrange_derivative(const std::vector<double> &xs, int j,
                                long double x)
{
  long double sum = 0.0L;
  for (size_t m = 0; m < xs.size(); ++m)
    {
      if ((int)m == j)
        continue;
      long double prod = 1.0L / ((long double)xs[j] - xs[m]);
      for (size_t l = 0; l < xs.size(); ++l)
        if ((int)l != j && l != m)
          prod *= (x - xs[l]) / ((long double)xs[j] - xs[l]);
      sum += prod;
    }
  return sum;
}

// Solve A X = B in place (A n x n, B n x m), Gaussian elimination with
// partial pivoting in long double.
void solve_dense(int n, int m, std::vector<long double> &A,
                 std::vector<long double> &B)
{
  for (int c = 0; c < n; ++c)
    {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (std::fabs((double)A[r * n + c]) > std::fabs((double)A[piv * n + c]))
          piv = r;
      if (piv != c)
        {
          for (int j = 0; j < n; ++j)
            std::swap(A[c * n + j], A[piv * n + j]);
          for (int j = 0; j < m; ++j)
            std::swap(B[c * m + j], B[piv * m + j]);
        }
      for (int r = 0; r < n; ++r)
        {
          if (r == c)
            continue;
          long double f = A[r * n + c] / A[c * n + c];
          for (int j = 0; j < n; ++j)
            A[r * n + j] -= f * A[c * n + j];
          for (int j = 0; j < m; ++j)
            B[r * m + j] -= f * B[c * m + j];
        }
    }
  for (int r = 0; r < n; ++r)
    for (int j = 0; j < m; ++j)
      B[r * m + j] /= A[r * n + r];
}

// The 1D operators the sum-factorised kernels use (hyper.deal gets them from
// deal.II's ShapeInfo: shape_values_eo, shape_gradients_collocation_eo,
// inverse_shape_values_eo, quadrature_data_on_face).
struct Tables1D
{
  int nd = 0, nq = 0;
  bool collocation = false;
  std::vector<double> xd, wd; // nodes: Gauss-Lobatto(k+1)
  std::vector<double> xq, wq; // quadrature: Gauss(nq) or Gauss-Lobatto(nq)
  std::vector<double> S;      // nq x nd: nodal basis at quadrature points
  std::vector<double> D;      // nq x nq: d/dx of the quadrature-point
                              //          Lagrange basis at quadrature points
  std::vector<double> F[2];   // nq: quadrature-point basis at 0 and 1
  std::vector<double> Sinv;   // nd x nq: (S^T W S)^{-1} S^T W

  void build(int degree, int n_points, bool coll)
  {
    nd = degree + 1;
    nq = n_points;
    collocation = coll;
    gauss_lobatto_rule(nd, xd, wd);
    if (coll)
      gauss_lobatto_rule(nq, xq, wq);
    else
      gauss_rule(nq, xq, wq);

    S.assign(nq * nd, 0.0);
    for (int q = 0; q < nq; ++q)
      for (int i = 0; i < nd; ++i)
        S[q * nd + i] = (double)lagrange_value(xd, i, xq[q]);
    D.assign(nq * nq, 0.0);
    for (int q = 0; q < nq; ++q)
      for (int p = 0; p < nq; ++p)
        D[q * nq + p] = (double)lagrange_derivative(xq, p, xq[q]);
    for (int s = 0; s < 2; ++s)
      {
        F[s].assign(nq, 0.0);
        for (int p = 0; p < nq; ++p)
          F[s][p] = (double)lagrange_value(xq, p, (long double)s);
      }
    // Sinv = (S^T W S)^{-1} S^T W
    std::vector<long double> M(nd * nd, 0.0L), R(nd * nq, 0.0L);
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j)
        for (int q = 0; q < nq; ++q)
          M[i * nd + j] +=
            (long double)S[q * nd + i] * wq[q] * (long double)S[q * nd + j];
    for (int i = 0; i < nd; ++i)
      for (int q = 0; q < nq; ++q)
        R[i * nq + q] = (long double)S[q * nd + i] * wq[q];
    solve_dense(nd, nq, M, R);
    Sinv.assign(nd * nq, 0.0);
    for (int i = 0; i < nd * nq; ++i)
      Sinv[i] = (double)R[i];

    if (coll)
      {
        // collocation: quadrature points are the nodes; S and Sinv are the
        // identity and the face vectors pick the end nodes
        for (int q = 0; q < nq; ++q)
          for (int i = 0; i < nd; ++i)
            {
              S[q * nd + i] = (q == i) ? 1.0 : 0.0;
              Sinv[i * nq + q] = (q == i) ? 1.0 : 0.0;
            }
        for (int p = 0; p < nq; ++p)
          {
            F[0][p] = (p == 0) ? 1.0 : 0.0;
            F[1][p] = (p == nq - 1) ? 1.0 : 0.0;
          }
      }
  }
};

// ---------------------------------------------------------------------------
// analytic functions
// ---------------------------------------------------------------------------
double eval_function(int id, const std::vector<double> &params, int dim_x,
                     int D, const double *p, double t)
{
  switch (id)
    {
      case HDO_FUNC_ZERO:
        return 0.0;
      case HDO_FUNC_HYPERRECTANGLE:
        {
          // examples/advection/cases/hyperrectangle.h:33-55 (wave number 2)
          double r = std::sin(2.0 * PI * (p[0] - t * params[0]));
          for (int d = 1; d < D; ++d)
            r *= std::cos(2.0 * PI * (p[d] - t * params[d]));
          return r;
        }
      case HDO_FUNC_CONSTANT_ONE:
        return 1.0;
      case HDO_FUNC_SIN_COS:
        {
          // tests/tests_functions.h:21-38
          double r = std::sin(p[0] * PI);
          for (int d = 1; d < D; ++d)
            r *= std::cos(p[d] * PI);
          return r;
        }
      case HDO_FUNC_PRODUCT_SINES:
        {
          double r = 1.0;
          for (int d = 0; d < D; ++d)
            r *= std::sin(2.0 * PI * (p[d] - t * params[d]));
          return r;
        }
      case HDO_FUNC_LANDAU:
        {
          // params: [eps, kx, np]: 1 + eps prod_{d < np} cos(kx x_d) times
          // a Maxwellian in v (np = 1 is examples/vlasov_poisson/cases/
          // hyperrectangle.h:44-58)
          const double eps = params.size() > 0 ? params[0] : 0.01;
          const double kx = params.size() > 1 ? params[1] : 0.5;
          const int np = params.size() > 2 ? (int)params[2] : dim_x;
          double px = 1.0;
          for (int d = 0; d < np; ++d)
            px *= std::cos(kx * p[d]);
          double v2 = 0.0;
          for (int d = dim_x; d < D; ++d)
            v2 += p[d] * p[d];
          const int dv = D - dim_x;
          return (1.0 + eps * px) * std::exp(-0.5 * v2) /
                 std::pow(2.0 * PI, 0.5 * dv);
        }
      default:
        return 0.0;
    }
}

// ---------------------------------------------------------------------------
// lane-batched tensor contraction
// ---------------------------------------------------------------------------
// out(.., r, ..) (+)= sum_c A(r, c) in(.., c, ..) along direction d, where
// A(r, c) = M[r * sr + c * sc]; ext[] are the input extents (lanes innermost),
// the output has extent m along d.
void contract(int D, const int *ext, int d, int m, const double *M, int sr,
              int sc, const double *in, double *out, bool add)
{
  int pre = 1, post = 1;
  for (int j = 0; j < d; ++j)
    pre *= ext[j];
  for (int j = d + 1; j < D; ++j)
    post *= ext[j];
  const int n = ext[d];
  const int blk = pre * L;
  for (int p = 0; p < post; ++p)
    for (int r = 0; r < m; ++r)
      {
        double *o = out + (size_t)(p * m + r) * blk;
        const double *ib = in + (size_t)p * n * blk;
        if (!add)
          for (int i = 0; i < blk; ++i)
            o[i] = 0.0;
        for (int c = 0; c < n; ++c)
          {
            const double a = M[r * sr + c * sc];
            const double *ii = ib + (size_t)c * blk;
#pragma omp simd
            for (int i = 0; i < blk; ++i)
              o[i] += a * ii[i];
          }
      }
}

int ipow(int b, int e)
{
  int r = 1;
  for (int i = 0; i < e; ++i)
    r *= b;
  return r;
}

// ---------------------------------------------------------------------------
// LSRK coefficients (Kennedy, Carpenter, Lewis 2000), the four schemes of
// base/time_integrators.templates.h:32-84.
// ---------------------------------------------------------------------------
bool rk_coefficients(const std::string &type, std::vector<double> &ai,
                     std::vector<double> &bi)
{
  if (type == "rk33")
    {
      bi = {0.245170287303492, 0.184896052186740, 0.569933660509768};
      ai = {0.755726351946097, 0.386954477304099};
    }
  else if (type == "rk45")
    {
      bi = {1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
            -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
            5198255086312. / 14908931495163.};
      ai = {970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
            2251764453980. / 15575788980749.,
            26877169314380. / 34165994151039.};
    }
  else if (type == "rk47")
    {
      bi = {0.0941840925477795334, 0.149683694803496998, 0.285204742060440058,
            -0.122201846148053668, 0.0605151571191401122, 0.345986987898399296,
            0.186627171718797670};
      const double c[6] = {0.241566650129646868, 0.0423866513027719953,
                           0.215602732678803776, 0.232328007537583987,
                           0.256223412574146438, 0.0978694102142697230};
      ai.resize(6);
      for (int i = 0; i < 6; ++i)
        ai[i] = c[i] + bi[i];
    }
  else if (type == "rk59")
    {
      bi = {2274579626619. / 23610510767302., 693987741272. / 12394497460941.,
            -347131529483. / 15096185902911., 1144057200723. / 32081666971178.,
            1562491064753. / 11797114684756., 13113619727965. / 44346030145118.,
            393957816125. / 7825732611452., 720647959663. / 6565743875477.,
            3559252274877. / 14424734981077.};
      ai = {1107026461565. / 5417078080134.,  38141181049399. / 41724347789894.,
            493273079041. / 11940823631197.,  1851571280403. / 6147804934346.,
            11782306865191. / 62590030070788., 9452544825720. / 13648368537481.,
            4435885630781. / 26285702406235., 2357909744247. / 11371140753790.};
    }
  else
    return false;
  return true;
}

using Op = std::function<void(const double *src, double *dst, double t)>;

// Two-register LSRK step (time_integrators.templates.h:91-184).  The first
// stage copies sol into Ti; every stage evaluates K = op(Ti), then
// sol += b dt K and Ti = sol_old + a dt K (no Ti update in the last stage).
void lsrk_step(const std::vector<double> &ai, const std::vector<double> &bi,
               size_t n, double *sol, double *ki, double *ti, double t,
               double dt, const Op &op)
{
  auto stage = [&](double time, double b, double a) {
    op(ti, ki, time);
    if (a == 0.0)
      {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i)
          sol[i] = sol[i] + b * ki[i];
      }
    else
      {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i)
          {
            const double k = ki[i], s = sol[i];
            sol[i] = s + b * k;
            ti[i] = s + a * k;
          }
      }
  };
  std::memcpy(ti, sol, n * sizeof(double));
  stage(t, bi[0] * dt, ai[0] * dt);
  double sum_previous_bi = 0.0;
  for (size_t s = 1; s < bi.size(); ++s)
    {
      const double c = sum_previous_bi + ai[s - 1];
      stage(t + c * dt, bi[s] * dt, s == bi.size() - 1 ? 0.0 : ai[s] * dt);
      sum_previous_bi += bi[s - 1];
    }
}

} // namespace

// ---------------------------------------------------------------------------
// problem
// ---------------------------------------------------------------------------
struct hdo_problem
{
  int dim_x = 0, dim_v = 0, D = 0, degree = 0, nd = 0, nq = 0;
  bool collocation = false;
  int n[6] = {1, 1, 1, 1, 1, 1};
  double lo[6] = {0}, hi[6] = {0}, h[6] = {1, 1, 1, 1, 1, 1};
  bool periodic[6] = {true, true, true, true, true, true};
  long long n_cells = 0, dofs_cell = 0;
  double skew = 0.0;
  double a[6] = {0};
  int bc_kind = HDO_BC_INHOMOGENEOUS;
  int func_id = HDO_FUNC_ZERO;
  std::vector<double> func_params;

  // Vlasov-Poisson: x-velocity = v at the v quadrature point, v-velocity =
  // grad phi at the x quadrature point (examples/vlasov_poisson/include/
  // velocity_field_view.h:111-175)
  bool vp = false;
  std::vector<double> efield; // n_x_cells x nq^dim_x x dim_x
  std::vector<int> qidx;      // nq^D x D multi-indices of quadrature points

  // deformed mesh (grid_generator.cc:110-230, DeformedCubeManifold): each
  // space s (0 = x, 1 = v) maps x^ -> x^ + deform[s] prod_j sin(freq[s] pi
  // (x^_j - lo_j) / (hi_j - lo_j)) (1, ..., 1); deform = 0: Cartesian
  bool deformed = false;
  double deform[2] = {0.0, 0.0};
  int freq[2] = {2, 2};

  Tables1D t;       // tables of the operator quadrature
  Tables1D tg;      // Gauss(k+1) tables (FCL inverse mass, quad no 1)
  std::vector<double> jxw;     // nq^D
  std::vector<double> jxwf[6]; // per direction d: nq^{D-1} (j != d)
  std::vector<double> jxw_g;   // (k+1)^D with the FCL mass-matrix rule

  void cell_index(long long c, int *idx) const
  {
    for (int d = 0; d < D; ++d)
      {
        idx[d] = (int)(c % n[d]);
        c /= n[d];
      }
  }
  long long cell_id(const int *idx) const
  {
    long long c = 0;
    for (int d = D - 1; d >= 0; --d)
      c = c * n[d] + idx[d];
    return c;
  }
};

namespace
{
// Per-thread scratch for one macro cell.
struct Work
{
  std::vector<double> u, uq, buf, data, tmp, tmp2, fm, fp, fd, fw;
  explicit Work(const hdo_problem &P)
  {
    const int nmax = std::max(P.nd, P.nq);
    const size_t vol = (size_t)ipow(nmax, P.D) * L;
    const size_t face = (size_t)ipow(nmax, P.D - 1) * L;
    u.resize(vol);
    uq.resize(vol);
    buf.resize(vol);
    data.resize(vol);
    tmp.resize(vol);
    tmp2.resize(vol);
    fm.resize(face);
    fp.resize(face);
    fd.resize(face);
    fw.resize(face);
  }
};

// Apply the 1D matrix along every direction in [d0, d1) except `skip`;
// ping-pongs between a and b and returns the buffer holding the result.
// ext is updated to the output extents.
double *contract_dirs(const hdo_problem &P, int *ext, int skip, int m,
                      const double *M, int sr, int sc, double *a, double *b,
                      bool reverse)
{
  double *cur = a, *nxt = b;
  for (int s = 0; s < P.D; ++s)
    {
      const int d = reverse ? P.D - 1 - s : s;
      if (d == skip)
        continue;
      contract(P.D, ext, d, m, M, sr, sc, cur, nxt, false);
      ext[d] = m;
      std::swap(cur, nxt);
    }
  return cur;
}

// Physical coordinates of the lane's face quadrature point.
void face_point(const hdo_problem &P, const int *idx, int d, int side,
                int fq, double *x)
{
  int rem = fq;
  for (int j = 0; j < P.D; ++j)
    {
      if (j == d)
        {
          x[j] = P.lo[j] + (idx[j] + side) * P.h[j];
          continue;
        }
      const int qj = rem % P.nq;
      rem /= P.nq;
      x[j] = P.lo[j] + (idx[j] + P.t.xq[qj]) * P.h[j];
    }
}

// Macro cell m: the L cell ids, their validity and multi-indices.
void macro_cells(const hdo_problem &P, long long m, long long *cells,
                 bool *valid, int idx[L][6])
{
  for (int l = 0; l < L; ++l)
    {
      const long long c = m * L + l;
      valid[l] = c < P.n_cells;
      cells[l] = valid[l] ? c : P.n_cells - 1;
      P.cell_index(cells[l], idx[l]);
    }
}

// Gather the nd^{D-1} nodal values of face (d, node_d) of each lane's cell
// (per-lane cell id in nb) into out with lanes innermost.
void gather_face_nodes(const hdo_problem &P, const double *src,
                       const long long *nb, int d, int node_d, double *out)
{
  const int nf = ipow(P.nd, P.D - 1);
  for (int f = 0; f < nf; ++f)
    {
      // expand face index (directions != d, lexicographic) to a cell node
      int rem = f, node = 0, stride = 1;
      for (int j = 0; j < P.D; ++j)
        {
          int ij;
          if (j == d)
            ij = node_d;
          else
            {
              ij = rem % P.nd;
              rem /= P.nd;
            }
          node += ij * stride;
          stride *= P.nd;
        }
      for (int l = 0; l < L; ++l)
        out[f * L + l] = src[nb[l] * P.dofs_cell + node];
    }
}

void scatter_add_face_nodes(const hdo_problem &P, double *dst,
                            const long long *cells, const bool *valid, int d,
                            int node_d, const double *in)
{
  const int nf = ipow(P.nd, P.D - 1);
  for (int f = 0; f < nf; ++f)
    {
      int rem = f, node = 0, stride = 1;
      for (int j = 0; j < P.D; ++j)
        {
          int ij;
          if (j == d)
            ij = node_d;
          else
            {
              ij = rem % P.nd;
              rem /= P.nd;
            }
          node += ij * stride;
          stride *= P.nd;
        }
      for (int l = 0; l < L; ++l)
        if (valid[l])
          dst[cells[l] * P.dofs_cell + node] += in[f * L + l];
    }
}

// ---------------------------------------------------------------------------
// Deformed phase-space geometry (evaluated per point, textbook formulas):
// the Jacobian J of the cell map xi -> x (block diagonal over the x and v
// spaces), its inverse by Gauss-Jordan elimination, det J, the physical
// normal J^{-T} n^ / |J^{-T} n^| and the face measure det J |J^{-T} n^|
// (the quantities FEEvaluation / FEFaceEvaluation take from deal.II's
// mapping: fe_evaluation_cell.h:196-271, fe_evaluation_face.h:434-463).
struct PointGeom
{
  double Ji[6][6]; // inverse Jacobian, Ji[i][j] = d xi_i / d x_j
  double det;
};

// physical point of the reference point x^ (per direction) of the mesh map
void map_point(const hdo_problem &P, const double *xh, double *x)
{
  for (int s = 0; s < 2; ++s)
    {
      const int b = s == 0 ? 0 : P.dim_x, e = s == 0 ? P.dim_x : P.D;
      double sv = P.deform[s];
      for (int j = b; j < e; ++j)
        sv *= std::sin(P.freq[s] * PI * (xh[j] - P.lo[j]) / (P.hi[j] - P.lo[j]));
      for (int j = b; j < e; ++j)
        x[j] = xh[j] + sv;
    }
}

void point_geometry(const hdo_problem &P, const int *idx, const double *xi,
                    PointGeom &g)
{
  const int D = P.D;
  double J[6][6] = {};
  for (int s = 0; s < 2; ++s)
    {
      const int b = s == 0 ? 0 : P.dim_x, e = s == 0 ? P.dim_x : P.D;
      double sn[6], cs[6];
      for (int j = b; j < e; ++j)
        {
          const double xh = P.lo[j] + (idx[j] + xi[j]) * P.h[j];
          const double arg = P.freq[s] * PI * (xh - P.lo[j]) / (P.hi[j] - P.lo[j]);
          sn[j] = std::sin(arg);
          cs[j] = std::cos(arg);
        }
      for (int j = b; j < e; ++j)
        {
          // d/dx^_j of deform prod_k sin(...)
          double ds = P.deform[s] * P.freq[s] * PI / (P.hi[j] - P.lo[j]) * cs[j];
          for (int k = b; k < e; ++k)
            if (k != j)
              ds *= sn[k];
          for (int i = b; i < e; ++i)
            J[i][j] = ((i == j ? 1.0 : 0.0) + ds) * P.h[j];
        }
    }
  // Gauss-Jordan with partial pivoting on [J | I]
  double A[6][12];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < 2 * D; ++j)
      A[i][j] = j < D ? J[i][j] : (j - D == i ? 1.0 : 0.0);
  double det = 1.0;
  for (int c = 0; c < D; ++c)
    {
      int piv = c;
      for (int r = c + 1; r < D; ++r)
        if (std::fabs(A[r][c]) > std::fabs(A[piv][c]))
          piv = r;
      if (piv != c)
        {
          for (int j = 0; j < 2 * D; ++j)
            std::swap(A[c][j], A[piv][j]);
          det = -det;
        }
      const double pv = A[c][c];
      det *= pv;
      for (int j = 0; j < 2 * D; ++j)
        A[c][j] /= pv;
      for (int r = 0; r < D; ++r)
        if (r != c && A[r][c] != 0.0)
          {
            const double f = A[r][c];
            for (int j = 0; j < 2 * D; ++j)
              A[r][j] -= f * A[c][j];
          }
    }
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      g.Ji[i][j] = A[i][D + j];
  g.det = det;
}

// JxW at the cell quadrature point with per-direction indices qi (points
// xq, weights wq of the given table)
double cell_jxw(const hdo_problem &P, const int *idx, const int *qi,
                const std::vector<double> &xq, const std::vector<double> &wq,
                PointGeom *out = nullptr)
{
  double xi[6] = {0, 0, 0, 0, 0, 0};
  double w = 1.0;
  for (int j = 0; j < P.D; ++j)
    {
      xi[j] = xq[qi[j]];
      w *= wq[qi[j]];
    }
  PointGeom g;
  point_geometry(P, idx, xi, g);
  if (out)
    *out = g;
  return std::fabs(g.det) * w;
}

// a . n (outward normal of side `side` of direction d) and the face JxW at
// face point fq of the cell idx
void face_geometry(const hdo_problem &P, const int *idx, int d, int side,
                   int fq, double &an, double &jxwf)
{
  double xi[6] = {0, 0, 0, 0, 0, 0};
  double w = 1.0;
  int rem = fq;
  for (int j = 0; j < P.D; ++j)
    {
      if (j == d)
        {
          xi[j] = side;
          continue;
        }
      const int qj = rem % P.nq;
      rem /= P.nq;
      xi[j] = P.t.xq[qj];
      w *= P.t.wq[qj];
    }
  PointGeom g;
  point_geometry(P, idx, xi, g);
  // J^{-T} n^ with n^ = +-e_d: component i is Ji[d][i]
  const double sg = side ? 1.0 : -1.0;
  double nn = 0.0;
  for (int i = 0; i < P.D; ++i)
    nn += g.Ji[d][i] * g.Ji[d][i];
  nn = std::sqrt(nn);
  an = 0.0;
  for (int i = 0; i < P.D; ++i)
    an += P.a[i] * sg * g.Ji[d][i] / nn;
  jxwf = std::fabs(g.det) * nn * w;
}

// Cell integral (a u, grad phi) in quadrature space, including the
// skew-symmetric part (advection_operation.h:318-403 / 713-787):
//   data = -skew (a . grad u) JxW + sum_d D_d^T [ (1-skew) u a_d / h_d JxW ]
// buf holds u at the quadrature points; tmp/tmp2 are scratch.
// Transport velocity of direction d at quadrature point qi of the cell idx
// (constant, or the phase-space field of the Vlasov-Poisson problem; for a
// face point qi[d] is not used).
inline double velocity(const hdo_problem &P, int d, const int *idx,
                       const int *qi)
{
  if (!P.vp)
    return P.a[d];
  if (d < P.dim_x)
    {
      const int e = P.dim_x + d;
      return P.lo[e] + (idx[e] + P.t.xq[qi[e]]) * P.h[e];
    }
  const int e = d - P.dim_x;
  long long xc = 0;
  int qx = 0;
  for (int j = P.dim_x - 1; j >= 0; --j)
    {
      xc = xc * P.n[j] + idx[j];
      qx = qx * P.nq + qi[j];
    }
  const int nqx = ipow(P.nq, P.dim_x);
  return P.efield[((size_t)xc * nqx + qx) * P.dim_x + e];
}

void cell_term(const hdo_problem &P, Work &w, const int (*idx)[6] = nullptr)
{
  const int D = P.D, nq = P.nq;
  const size_t nvol = (size_t)ipow(nq, D);
  int ext[6];
  for (int j = 0; j < D; ++j)
    ext[j] = nq;
  std::fill(w.data.begin(), w.data.begin() + nvol * L, 0.0);
  if (P.deformed)
    {
      // a . grad v = (J^{-1} a) . grad^ v at every point:
      //   -skew (a . grad u) JxW + sum_d D_d^T [(1-skew) u (J^{-1} a)_d JxW]
      std::vector<double> ah(nvol * L * D), jw(nvol * L);
      for (size_t q = 0; q < nvol; ++q)
        for (int l = 0; l < L; ++l)
          {
            PointGeom g;
            jw[q * L + l] = cell_jxw(P, idx[l], &P.qidx[q * D], P.t.xq,
                                     P.t.wq, &g);
            for (int d = 0; d < D; ++d)
              {
                double v = 0.0;
                for (int i = 0; i < D; ++i)
                  v += g.Ji[d][i] * P.a[i];
                ah[(q * L + l) * D + d] = v;
              }
          }
      for (int d = 0; d < D; ++d)
        {
          if (P.skew != 0.0)
            contract(D, ext, d, nq, P.t.D.data(), nq, 1, w.buf.data(),
                     w.tmp2.data(), false);
          for (size_t q = 0; q < nvol; ++q)
            for (int l = 0; l < L; ++l)
              {
                const double vel = ah[(q * L + l) * D + d] * jw[q * L + l];
                if (P.skew != 0.0)
                  w.data[q * L + l] += -P.skew * w.tmp2[q * L + l] * vel;
                w.tmp[q * L + l] = (1.0 - P.skew) * w.buf[q * L + l] * vel;
              }
          if (P.skew != 1.0)
            contract(D, ext, d, nq, P.t.D.data(), 1, nq, w.tmp.data(),
                     w.data.data(), true);
        }
      return;
    }
  if (P.vp)
    {
      // per-point velocity: -skew (a . grad u) JxW + D_d^T[(1-skew) u a_d
      // / h_d JxW] with a = a(x, v)
      for (int d = 0; d < D; ++d)
        {
          if (P.skew != 0.0)
            contract(D, ext, d, nq, P.t.D.data(), nq, 1, w.buf.data(),
                     w.tmp2.data(), false);
          for (size_t q = 0; q < nvol; ++q)
            for (int l = 0; l < L; ++l)
              {
                const double vel =
                  velocity(P, d, idx[l], &P.qidx[q * D]) / P.h[d] * P.jxw[q];
                if (P.skew != 0.0)
                  w.data[q * L + l] += -P.skew * w.tmp2[q * L + l] * vel;
                w.tmp[q * L + l] = (1.0 - P.skew) * w.buf[q * L + l] * vel;
              }
          if (P.skew != 1.0)
            contract(D, ext, d, nq, P.t.D.data(), 1, nq, w.tmp.data(),
                     w.data.data(), true);
        }
      return;
    }
  if (P.skew != 0.0)
    for (int d = 0; d < D; ++d)
      {
        contract(D, ext, d, nq, P.t.D.data(), nq, 1, w.buf.data(),
                 w.tmp.data(), false);
        const double c = -P.skew * P.a[d] / P.h[d];
        for (size_t q = 0; q < nvol; ++q)
          {
            const double s = c * P.jxw[q];
            for (int l = 0; l < L; ++l)
              w.data[q * L + l] += w.tmp[q * L + l] * s;
          }
      }
  if (P.skew != 1.0)
    for (int d = 0; d < D; ++d)
      {
        const double c = (1.0 - P.skew) * P.a[d] / P.h[d];
        for (size_t q = 0; q < nvol; ++q)
          {
            const double s = c * P.jxw[q];
            for (int l = 0; l < L; ++l)
              w.tmp[q * L + l] = w.buf[q * L + l] * s;
          }
        // transposed collocation gradient: A(r, c) = D[c][r]
        contract(D, ext, d, nq, P.t.D.data(), 1, nq, w.tmp.data(),
                 w.data.data(), true);
      }
}

// Upwind numerical flux tested from the minus side, already multiplied by
// -JxW_face (fe_evaluation_face.h:381-399 and advection_operation.h:455-519).
inline double face_value(double um, double up, double an, double skew,
                         double jxwf)
{
  const double flux = 0.5 * ((um + up) * an + std::fabs(an) * (um - up));
  return -(flux - skew * um * an) * jxwf;
}

// ECL: one macro cell of local_apply_advect_and_inverse_mass_matrix
// (advection_operation.h:221-566).
void ecl_cells(const hdo_problem &P, Work &w, const double *src, double *dst,
               double t, const long long *cells, const bool *valid);

void ecl_macro(const hdo_problem &P, Work &w, const double *src, double *dst,
               double t, long long m)
{
  long long cells[L];
  bool valid[L];
  int idx[L][6];
  macro_cells(P, m, cells, valid, idx);
  ecl_cells(P, w, src, dst, t, cells, valid);
}

// ECL on the L cells given (lanes with valid == false are computed but not
// written)
void ecl_cells(const hdo_problem &P, Work &w, const double *src, double *dst,
               double t, const long long *cells, const bool *valid)
{
  const int D = P.D, nd = P.nd, nq = P.nq;
  int idx[L][6];
  for (int l = 0; l < L; ++l)
    P.cell_index(cells[l], idx[l]);

  // 1) read and interpolate to the quadrature points
  const size_t ndofs = (size_t)P.dofs_cell;
  for (size_t i = 0; i < ndofs; ++i)
    for (int l = 0; l < L; ++l)
      w.u[i * L + l] = src[cells[l] * ndofs + i];
  int ext[6];
  for (int j = 0; j < D; ++j)
    ext[j] = nd;
  double *uq = w.u.data();
  if (!P.collocation)
    uq = contract_dirs(P, ext, -1, nq, P.t.S.data(), nd, 1, w.u.data(),
                       w.uq.data(), false);
  const size_t nvol = (size_t)ipow(nq, D);
  std::copy(uq, uq + nvol * L, w.buf.begin());

  // 2) cell integral
  cell_term(P, w, idx);

  // 3) faces
  const int nfq = ipow(nq, D - 1);
  for (int f = 0; f < 2 * D; ++f)
    {
      const int d = f / 2, side = f % 2;
      const double nrm = side ? 1.0 : -1.0;
      int fext[6];
      for (int j = 0; j < D; ++j)
        fext[j] = nq;
      // u^- : trace of the quadrature-point interpolant (A = F[side], 1 x nq)
      contract(D, fext, d, 1, P.t.F[side].data(), 0, 1, w.buf.data(),
               w.fm.data(), false);

      // u^+ : the neighbour's face nodes, interpolated along the face
      long long nb[L];
      bool boundary[L];
      for (int l = 0; l < L; ++l)
        {
          int j[6];
          std::copy(idx[l], idx[l] + D, j);
          j[d] += side ? 1 : -1;
          boundary[l] = false;
          if (j[d] < 0 || j[d] >= P.n[d])
            {
              if (P.periodic[d])
                j[d] = (j[d] + P.n[d]) % P.n[d];
              else
                {
                  boundary[l] = true;
                  j[d] = idx[l][d];
                }
            }
          nb[l] = P.cell_id(j);
        }
      gather_face_nodes(P, src, nb, d, side ? 0 : nd - 1, w.fw.data());
      int gext[6];
      for (int j = 0; j < D; ++j)
        gext[j] = nd;
      gext[d] = 1;
      double *up = w.fw.data();
      if (!P.collocation)
        up = contract_dirs(P, gext, d, nq, P.t.S.data(), nd, 1, w.fw.data(),
                           w.fp.data(), false);

      for (int l = 0; l < L; ++l)
        if (boundary[l])
          for (int q = 0; q < nfq; ++q)
            {
              double g = 0.0;
              if (P.bc_kind == HDO_BC_INHOMOGENEOUS)
                {
                  double xh[6], x[6];
                  face_point(P, idx[l], d, side, q, xh);
                  map_point(P, xh, x);
                  g = eval_function(P.func_id, P.func_params, P.dim_x, D, x,
                                    t);
                }
              up[q * L + l] = -w.fm[q * L + l] + 2.0 * g;
            }

      if (P.deformed)
        {
          for (int q = 0; q < nfq; ++q)
            for (int l = 0; l < L; ++l)
              {
                double an, jf;
                face_geometry(P, idx[l], d, side, q, an, jf);
                w.fd[q * L + l] =
                  face_value(w.fm[q * L + l], up[q * L + l], an, P.skew, jf);
              }
        }
      else if (!P.vp)
        {
          const double an = P.a[d] * nrm;
          for (int q = 0; q < nfq; ++q)
            for (int l = 0; l < L; ++l)
              w.fd[q * L + l] = face_value(w.fm[q * L + l], up[q * L + l],
                                           an, P.skew, P.jxwf[d][q]);
        }
      else
        for (int q = 0; q < nfq; ++q)
          {
            int qi[6] = {0, 0, 0, 0, 0, 0};
            int rem = q;
            for (int j = 0; j < D; ++j)
              if (j != d)
                {
                  qi[j] = rem % nq;
                  rem /= nq;
                }
            for (int l = 0; l < L; ++l)
              {
                const double an = velocity(P, d, idx[l], qi) * nrm;
                w.fd[q * L + l] = face_value(w.fm[q * L + l], up[q * L + l],
                                             an, P.skew, P.jxwf[d][q]);
              }
          }
      // lift: data(.., p, ..) += F[side][p] * fd  (A(r, 0) = F[side][r])
      fext[d] = 1;
      contract(D, fext, d, nq, P.t.F[side].data(), 1, 0, w.fd.data(),
               w.data.data(), true);
    }

  // 4) inverse mass matrix: divide by JxW, then Sinv along each direction
  for (size_t q = 0; q < nvol; ++q)
    for (int l = 0; l < L; ++l)
      w.data[q * L + l] /=
        P.deformed ? cell_jxw(P, idx[l], &P.qidx[q * D], P.t.xq, P.t.wq)
                   : P.jxw[q];
  for (int j = 0; j < D; ++j)
    ext[j] = nq;
  double *res = w.data.data();
  if (!P.collocation)
    res = contract_dirs(P, ext, -1, nd, P.t.Sinv.data(), nq, 1, w.data.data(),
                        w.tmp.data(), true);
  for (int l = 0; l < L; ++l)
    if (valid[l])
      for (size_t i = 0; i < ndofs; ++i)
        dst[cells[l] * ndofs + i] = res[i * L + l];
}

// FCL cell term (advection_operation.h:639-818): the cell integral tested
// with the nodal basis, written (not added) into dst.
void fcl_cell_macro(const hdo_problem &P, Work &w, const double *src,
                    double *dst, long long m)
{
  const int D = P.D, nd = P.nd, nq = P.nq;
  long long cells[L];
  bool valid[L];
  int idx[L][6];
  macro_cells(P, m, cells, valid, idx);
  const size_t ndofs = (size_t)P.dofs_cell;
  for (size_t i = 0; i < ndofs; ++i)
    for (int l = 0; l < L; ++l)
      w.u[i * L + l] = src[cells[l] * ndofs + i];
  int ext[6];
  for (int j = 0; j < D; ++j)
    ext[j] = nd;
  double *uq = w.u.data();
  if (!P.collocation)
    uq = contract_dirs(P, ext, -1, nq, P.t.S.data(), nd, 1, w.u.data(),
                       w.uq.data(), false);
  const size_t nvol = (size_t)ipow(nq, D);
  std::copy(uq, uq + nvol * L, w.buf.begin());
  cell_term(P, w, idx);
  for (int j = 0; j < D; ++j)
    ext[j] = nq;
  double *res = w.data.data();
  if (!P.collocation) // S^T along each direction: A(r, c) = S[c][r]
    res = contract_dirs(P, ext, -1, nd, P.t.S.data(), 1, nd, w.data.data(),
                        w.tmp.data(), true);
  for (int l = 0; l < L; ++l)
    if (valid[l])
      for (size_t i = 0; i < ndofs; ++i)
        dst[cells[l] * ndofs + i] = res[i * L + l];
}

// FCL faces in direction d for the macro cell m of minus-side cells
// (advection_operation.h:825-1131): interior faces add to both cells,
// boundary faces only to the cell inside.
void fcl_faces_macro(const hdo_problem &P, Work &w, const double *src,
                     double *dst, double t, int d, long long m)
{
  const int D = P.D, nd = P.nd, nq = P.nq;
  long long cells[L];
  bool valid[L];
  int idx[L][6];
  macro_cells(P, m, cells, valid, idx);
  const int nfq = ipow(nq, D - 1);
  int gext[6];

  // interior face between cell (minus) and its right neighbour (plus)
  long long plus[L];
  bool vin[L], vplus[L];
  for (int l = 0; l < L; ++l)
    {
      int j[6];
      std::copy(idx[l], idx[l] + D, j);
      j[d] += 1;
      vin[l] = valid[l];
      if (j[d] >= P.n[d])
        {
          if (P.periodic[d])
            j[d] = 0;
          else
            vin[l] = false;
        }
      plus[l] = P.cell_id(j[d] >= P.n[d] ? idx[l] : j);
      vplus[l] = vin[l];
    }
  auto to_quad = [&](double *nodal, double *alt) -> double * {
    for (int j = 0; j < D; ++j)
      gext[j] = nd;
    gext[d] = 1;
    if (P.collocation)
      return nodal;
    return contract_dirs(P, gext, d, nq, P.t.S.data(), nd, 1, nodal, alt,
                         false);
  };
  auto to_nodal = [&](double *q, double *alt) -> double * {
    for (int j = 0; j < D; ++j)
      gext[j] = nq;
    gext[d] = 1;
    if (P.collocation)
      return q;
    return contract_dirs(P, gext, d, nd, P.t.S.data(), 1, nd, q, alt, true);
  };

  bool any = false;
  for (int l = 0; l < L; ++l)
    any = any || vin[l];
  if (any)
    {
      gather_face_nodes(P, src, cells, d, nd - 1, w.fw.data());
      double *um = to_quad(w.fw.data(), w.fm.data());
      if (um != w.fm.data())
        std::copy(um, um + (size_t)nfq * L, w.fm.begin());
      gather_face_nodes(P, src, plus, d, 0, w.fw.data());
      double *up = to_quad(w.fw.data(), w.fp.data());
      if (up != w.fp.data())
        std::copy(up, up + (size_t)nfq * L, w.fp.begin());
      for (int q = 0; q < nfq; ++q)
        for (int l = 0; l < L; ++l)
          {
            // the normal of the minus side (+e_d on the reference cell)
            double an = P.a[d], jf = P.jxwf[d][q];
            if (P.deformed)
              face_geometry(P, idx[l], d, 1, q, an, jf);
            const double u_m = w.fm[q * L + l], u_p = w.fp[q * L + l];
            const double flux =
              0.5 * ((u_m + u_p) * an + std::fabs(an) * (u_m - u_p));
            w.fm[q * L + l] = -(flux - P.skew * u_m * an) * jf;
            w.fp[q * L + l] = -(-flux + P.skew * u_p * an) * jf;
          }
      double *rm = to_nodal(w.fm.data(), w.fd.data());
      scatter_add_face_nodes(P, dst, cells, vin, d, nd - 1, rm);
      double *rp = to_nodal(w.fp.data(), w.fd.data());
      scatter_add_face_nodes(P, dst, plus, vplus, d, 0, rp);
    }

  // boundary faces of non-periodic directions
  if (!P.periodic[d])
    for (int side = 0; side < 2; ++side)
      {
        bool vb[L];
        bool anyb = false;
        for (int l = 0; l < L; ++l)
          {
            vb[l] = valid[l] && (side ? idx[l][d] == P.n[d] - 1
                                      : idx[l][d] == 0);
            anyb = anyb || vb[l];
          }
        if (!anyb)
          continue;
        gather_face_nodes(P, src, cells, d, side ? nd - 1 : 0, w.fw.data());
        double *um = to_quad(w.fw.data(), w.fm.data());
        if (um != w.fm.data())
          std::copy(um, um + (size_t)nfq * L, w.fm.begin());
        for (int q = 0; q < nfq; ++q)
          for (int l = 0; l < L; ++l)
            {
              double an = P.a[d] * (side ? 1.0 : -1.0), jf = P.jxwf[d][q];
              if (P.deformed)
                face_geometry(P, idx[l], d, side, q, an, jf);
              double g = 0.0;
              if (P.bc_kind == HDO_BC_INHOMOGENEOUS && vb[l])
                {
                  double xh[6], x[6];
                  face_point(P, idx[l], d, side, q, xh);
                  map_point(P, xh, x);
                  g = eval_function(P.func_id, P.func_params, P.dim_x, D, x,
                                    t);
                }
              const double u_m = w.fm[q * L + l];
              const double u_p = -u_m + 2.0 * g;
              w.fd[q * L + l] = face_value(u_m, u_p, an, P.skew, jf);
            }
        double *r = to_nodal(w.fd.data(), w.fw.data());
        scatter_add_face_nodes(P, dst, cells, vb, d, side ? nd - 1 : 0, r);
      }
}

// FCL inverse mass (advection_operation.h:573-632) with the Gauss(k+1) rule
// (collocation: the Gauss-Lobatto rule): S^{-T}, divide by JxW, S^{-1}.
void fcl_mass_macro(const hdo_problem &P, Work &w, double *dst, long long m)
{
  const int D = P.D, nd = P.nd;
  long long cells[L];
  bool valid[L];
  int idx[L][6];
  macro_cells(P, m, cells, valid, idx);
  const size_t ndofs = (size_t)P.dofs_cell;
  for (size_t i = 0; i < ndofs; ++i)
    for (int l = 0; l < L; ++l)
      w.u[i * L + l] = dst[cells[l] * ndofs + i];
  const Tables1D &tg = P.collocation ? P.t : P.tg;
  int ext[6];
  for (int j = 0; j < D; ++j)
    ext[j] = nd;
  double *cur = w.u.data();
  if (!P.collocation) // Sinv^T: A(r, c) = Sinv[c][r]
    cur = contract_dirs(P, ext, -1, nd, tg.Sinv.data(), 1, nd, w.u.data(),
                        w.uq.data(), false);
  for (size_t q = 0; q < ndofs; ++q)
    {
      int qi[6];
      int rem = (int)q;
      for (int j = 0; j < D; ++j)
        {
          qi[j] = rem % nd;
          rem /= nd;
        }
      for (int l = 0; l < L; ++l)
        cur[q * L + l] /=
          P.deformed ? cell_jxw(P, idx[l], qi, tg.xq, tg.wq) : P.jxw_g[q];
    }
  double *other = (cur == w.u.data()) ? w.uq.data() : w.u.data();
  if (!P.collocation)
    cur = contract_dirs(P, ext, -1, nd, tg.Sinv.data(), nd, 1, cur, other,
                        true);
  for (int l = 0; l < L; ++l)
    if (valid[l])
      for (size_t i = 0; i < ndofs; ++i)
        dst[cells[l] * ndofs + i] = cur[i * L + l];
}

long long n_macro(const hdo_problem &P) { return (P.n_cells + L - 1) / L; }

void apply_ecl(const hdo_problem &P, const double *src, double *dst, double t)
{
  const long long nm = n_macro(P);
#pragma omp parallel
  {
    Work w(P);
#pragma omp for schedule(static)
    for (long long m = 0; m < nm; ++m)
      ecl_macro(P, w, src, dst, t, m);
  }
}

} // namespace

// FCL apply: cell phase (parallel), face phase (serial sweep per direction,
// since a face writes into two cells), inverse-mass phase (parallel).
static void apply_fcl(const hdo_problem &P, const double *src,
                           double *dst, double t)
{
  const long long nm = n_macro(P);
#pragma omp parallel
  {
    Work w(P);
#pragma omp for schedule(static)
    for (long long m = 0; m < nm; ++m)
      fcl_cell_macro(P, w, src, dst, m);
  }
  {
    Work w(P);
    for (int d = 0; d < P.D; ++d)
      for (long long m = 0; m < nm; ++m)
        fcl_faces_macro(P, w, src, dst, t, d, m);
  }
#pragma omp parallel
  {
    Work w(P);
#pragma omp for schedule(static)
    for (long long m = 0; m < nm; ++m)
      fcl_mass_macro(P, w, dst, m);
  }
}

// ---------------------------------------------------------------------------
// Vlasov-Poisson pieces (examples/vlasov_poisson)
// ---------------------------------------------------------------------------
namespace
{
struct XSpace
{
  int dx, nd, nq;
  long long n_cells, n_dofs;
  int dofs_cell;
  int n[3];
  double h[3];
  // exact 1D matrices of the nodal (Gauss-Lobatto) basis on [0,1]
  std::vector<double> M, K;      // nd x nd: int l_i l_j, int l_i' l_j'
  std::vector<double> d0, d1;    // l_i'(0), l_i'(1)
  std::vector<double> S, G, w;   // nq x nd values / derivatives at Gauss
                                 // points of the operator rule, weights
};

XSpace make_xspace(const hdo_problem &P)
{
  XSpace X;
  X.dx = P.dim_x;
  X.nd = P.nd;
  X.nq = P.nq;
  X.n_cells = 1;
  for (int d = 0; d < X.dx; ++d)
    {
      X.n[d] = P.n[d];
      X.h[d] = P.h[d];
      X.n_cells *= P.n[d];
    }
  X.dofs_cell = ipow(X.nd, X.dx);
  X.n_dofs = X.n_cells * X.dofs_cell;
  const auto &xd = P.t.xd;
  // exact mass / stiffness with Gauss(k+1)
  const auto &g = P.tg;
  X.M.assign(X.nd * X.nd, 0.0);
  X.K.assign(X.nd * X.nd, 0.0);
  for (int q = 0; q < g.nq; ++q)
    for (int i = 0; i < X.nd; ++i)
      for (int j = 0; j < X.nd; ++j)
        {
          const long double li = lagrange_value(xd, i, g.xq[q]);
          const long double lj = lagrange_value(xd, j, g.xq[q]);
          const long double di = lagrange_derivative(xd, i, g.xq[q]);
          const long double dj = lagrange_derivative(xd, j, g.xq[q]);
          X.M[i * X.nd + j] += (double)(g.wq[q] * li * lj);
          X.K[i * X.nd + j] += (double)(g.wq[q] * di * dj);
        }
  X.d0.resize(X.nd);
  X.d1.resize(X.nd);
  for (int i = 0; i < X.nd; ++i)
    {
      X.d0[i] = (double)lagrange_derivative(xd, i, 0.0L);
      X.d1[i] = (double)lagrange_derivative(xd, i, 1.0L);
    }
  X.S.assign(X.nq * X.nd, 0.0);
  X.G.assign(X.nq * X.nd, 0.0);
  X.w = P.t.wq;
  for (int q = 0; q < X.nq; ++q)
    for (int i = 0; i < X.nd; ++i)
      {
        X.S[q * X.nd + i] = (double)lagrange_value(xd, i, P.t.xq[q]);
        X.G[q * X.nd + i] = (double)lagrange_derivative(xd, i, P.t.xq[q]);
      }
  return X;
}

// y = (A_1 (x) ... (x) A_dx) x on one cell, A_j (m x nd) row-major
void tensor_apply(int dx, int nd, const std::vector<const double *> &A,
                  const std::vector<int> &m, const double *x, double *y)
{
  std::vector<double> cur(x, x + ipow(nd, dx)), nxt;
  std::vector<int> ext(dx, nd);
  for (int d = 0; d < dx; ++d)
    {
      int pre = 1, post = 1;
      for (int j = 0; j < d; ++j)
        pre *= ext[j];
      for (int j = d + 1; j < dx; ++j)
        post *= ext[j];
      nxt.assign((size_t)pre * m[d] * post, 0.0);
      for (int po = 0; po < post; ++po)
        for (int r = 0; r < m[d]; ++r)
          for (int c = 0; c < ext[d]; ++c)
            {
              const double a = A[d][r * ext[d] + c];
              for (int pi = 0; pi < pre; ++pi)
                nxt[pi + pre * (r + m[d] * po)] +=
                  a * cur[pi + pre * (c + ext[d] * po)];
            }
      ext[d] = m[d];
      cur.swap(nxt);
    }
  std::copy(cur.begin(), cur.end(), y);
}

void x_cell_index(const XSpace &X, long long c, int *idx)
{
  for (int d = 0; d < X.dx; ++d)
    {
      idx[d] = (int)(c % X.n[d]);
      c /= X.n[d];
    }
}
long long x_cell_id(const XSpace &X, const int *idx)
{
  long long c = 0;
  for (int d = X.dx - 1; d >= 0; --d)
    c = c * X.n[d] + idx[d];
  return c;
}

// Periodic SIPG Laplacian (examples/vlasov_poisson/include/poisson.h:100-254):
//   a(u, v) = sum_K (grad u, grad v) + sum_F ( sigma [u][v] - {d_n u}[v]
//             - [u]{d_n v} ),  sigma = k(k+1) (1/h- + 1/h+)
// with every integral exact (the reference's Gauss(k+1) rule is exact for
// these integrands), applied to src minus its mean (vmult, :100-123).
void sipg_apply(const XSpace &X, int degree, const double *src_in, double *dst)
{
  const long long nx = X.n_dofs;
  const int nd = X.nd, dc = X.dofs_cell, dx = X.dx;
  double mean = 0.0;
  for (long long i = 0; i < nx; ++i)
    mean += src_in[i];
  mean /= (double)nx;
  std::vector<double> src(src_in, src_in + nx);
  for (auto &v : src)
    v -= mean;
  std::fill(dst, dst + nx, 0.0);
  const double kk = std::max(degree, 1) * (degree + 1.0);
  // scaled 1D matrices per direction
  std::vector<std::vector<double>> Mh(dx), Kh(dx);
  for (int d = 0; d < dx; ++d)
    {
      Mh[d] = X.M;
      Kh[d] = X.K;
      for (auto &v : Mh[d])
        v *= X.h[d];
      for (auto &v : Kh[d])
        v /= X.h[d];
    }
  std::vector<double> tmp(dc);
  for (long long c = 0; c < X.n_cells; ++c)
    {
      // cell term: sum_d (x)_{j != d} h_j Mhat (x) Khat / h_d
      for (int d = 0; d < dx; ++d)
        {
          std::vector<const double *> A(dx);
          std::vector<int> m(dx, nd);
          for (int j = 0; j < dx; ++j)
            A[j] = j == d ? Kh[j].data() : Mh[j].data();
          tensor_apply(dx, nd, A, m, src.data() + c * dc, tmp.data());
          for (int i = 0; i < dc; ++i)
            dst[c * dc + i] += tmp[i];
        }
    }
  // faces: between cell L and its right neighbour R in direction d
  const int nf = ipow(nd, dx - 1);
  for (int d = 0; d < dx; ++d)
    {
      const double h = X.h[d];
      const double sigma = kk * (1.0 / h + 1.0 / h);
      const int sd = ipow(nd, d);
      for (long long c = 0; c < X.n_cells; ++c)
        {
          int idx[3];
          x_cell_index(X, c, idx);
          int jr[3] = {idx[0], idx[1], idx[2]};
          jr[d] = (idx[d] + 1) % X.n[d];
          const long long r = x_cell_id(X, jr);
          // 1D face quantities for every tangential node, then the
          // tangential mass
          std::vector<double> fl(nf * nd, 0.0), fr(nf * nd, 0.0);
          for (int f = 0; f < nf; ++f)
            {
              int rem = f, base = 0, pw = 1;
              for (int j = 0; j < dx; ++j)
                {
                  if (j != d)
                    {
                      base += (rem % nd) * pw;
                      rem /= nd;
                    }
                  pw *= nd;
                }
              const double *uL = src.data() + c * dc + base;
              const double *uR = src.data() + r * dc + base;
              const double um = uL[(nd - 1) * sd], upl = uR[0];
              double dm = 0.0, dp = 0.0;
              for (int i = 0; i < nd; ++i)
                {
                  dm += X.d1[i] * uL[i * sd] / h;
                  dp += X.d0[i] * uR[i * sd] / h;
                }
              const double jump = um - upl, avgd = 0.5 * (dm + dp);
              const double val = sigma * jump - avgd;
              for (int i = 0; i < nd; ++i)
                {
                  const double eL = (i == nd - 1) ? 1.0 : 0.0;
                  const double eR = (i == 0) ? 1.0 : 0.0;
                  fl[f * nd + i] = val * eL - 0.5 * jump * X.d1[i] / h;
                  fr[f * nd + i] = -val * eR - 0.5 * jump * X.d0[i] / h;
                }
            }
          // tangential mass on the face arrays (face index f over j != d,
          // normal index i), scatter into the cells
          for (int side = 0; side < 2; ++side)
            {
              const std::vector<double> &fa = side ? fr : fl;
              std::vector<double> out(nf * nd, 0.0);
              for (int i = 0; i < nd; ++i)
                {
                  std::vector<double> col(nf), res(nf);
                  for (int f = 0; f < nf; ++f)
                    col[f] = fa[f * nd + i];
                  if (dx == 1)
                    res = col;
                  else
                    {
                      std::vector<const double *> A;
                      std::vector<int> m;
                      for (int j = 0; j < dx; ++j)
                        if (j != d)
                          {
                            A.push_back(Mh[j].data());
                            m.push_back(nd);
                          }
                      tensor_apply(dx - 1, nd, A, m, col.data(), res.data());
                    }
                  for (int f = 0; f < nf; ++f)
                    out[f * nd + i] = res[f];
                }
              const long long cell = side ? r : c;
              for (int f = 0; f < nf; ++f)
                {
                  int rem = f, base = 0, pw = 1;
                  for (int j = 0; j < dx; ++j)
                    {
                      if (j != d)
                        {
                          base += (rem % nd) * pw;
                          rem /= nd;
                        }
                      pw *= nd;
                    }
                  for (int i = 0; i < nd; ++i)
                    dst[cell * dc + base + i * sd] += out[f * nd + i];
                }
            }
        }
    }
}

// rho(x node) = sum over v cells and v nodes of f JxW_v at the Gauss-Lobatto
// points (numerics/vector_tools.h:230-315)
void vp_density(const hdo_problem &P, const XSpace &X, const double *f,
                double *rho)
{
  std::fill(rho, rho + X.n_dofs, 0.0);
  const int dv = P.dim_v, nd = P.nd;
  const int dcx = X.dofs_cell, dcv = ipow(nd, dv);
  std::vector<double> wv(dcv);
  for (int q = 0; q < dcv; ++q)
    {
      double v = 1.0;
      int rem = q;
      for (int j = 0; j < dv; ++j)
        {
          v *= P.h[P.dim_x + j] * P.t.wd[rem % nd];
          rem /= nd;
        }
      wv[q] = v;
    }
  for (long long c = 0; c < P.n_cells; ++c)
    {
      const long long xc = c % X.n_cells;
      for (int qv = 0; qv < dcv; ++qv)
        for (int qx = 0; qx < dcx; ++qx)
          rho[xc * dcx + qx] += f[c * P.dofs_cell + qx + dcx * qv] * wv[qv];
    }
}

// b = -M (rho - mean), then b -= mean (vp/application.h:531-568)
void vp_rhs(const XSpace &X, double *rho, double *b)
{
  const long long nx = X.n_dofs;
  double mean = 0.0;
  for (long long i = 0; i < nx; ++i)
    mean += rho[i];
  mean /= (double)nx;
  for (long long i = 0; i < nx; ++i)
    rho[i] -= mean;
  std::vector<std::vector<double>> Mh(X.dx);
  std::vector<const double *> A(X.dx);
  std::vector<int> m(X.dx, X.nd);
  for (int d = 0; d < X.dx; ++d)
    {
      Mh[d] = X.M;
      for (auto &v : Mh[d])
        v *= X.h[d];
      A[d] = Mh[d].data();
    }
  for (long long c = 0; c < X.n_cells; ++c)
    tensor_apply(X.dx, X.nd, A, m, rho + c * X.dofs_cell, b + c * X.dofs_cell);
  for (long long i = 0; i < nx; ++i)
    b[i] = -b[i];
  mean = 0.0;
  for (long long i = 0; i < nx; ++i)
    mean += b[i];
  mean /= (double)nx;
  for (long long i = 0; i < nx; ++i)
    b[i] -= mean;
}

// conjugate gradients on the SIPG operator, warm-started from phi, until
// the relative residual ||b - A phi|| <= rel_tol ||b||; returns the
// iteration count
int poisson_cg(const XSpace &X, int degree, const double *b, double *phi,
               double rel_tol, int max_it)
{
  const long long n = X.n_dofs;
  std::vector<double> r(n), p(n), Ap(n);
  sipg_apply(X, degree, phi, Ap.data());
  double rr = 0.0, bb = 0.0;
  for (long long i = 0; i < n; ++i)
    {
      r[i] = b[i] - Ap[i];
      p[i] = r[i];
      rr += r[i] * r[i];
      bb += b[i] * b[i];
    }
  const double stop = rel_tol * std::sqrt(bb);
  int it = 0;
  while (std::sqrt(rr) > stop && it < max_it)
    {
      sipg_apply(X, degree, p.data(), Ap.data());
      double pAp = 0.0;
      for (long long i = 0; i < n; ++i)
        pAp += p[i] * Ap[i];
      if (!(pAp > 0.0))
        break; // breakdown: the residual left lies in the null space
      const double alpha = rr / pAp;
      double rr2 = 0.0;
      for (long long i = 0; i < n; ++i)
        {
          phi[i] += alpha * p[i];
          r[i] -= alpha * Ap[i];
          rr2 += r[i] * r[i];
        }
      const double beta = rr2 / rr;
      rr = rr2;
      for (long long i = 0; i < n; ++i)
        p[i] = r[i] + beta * p[i];
      ++it;
    }
  return it;
}

// grad phi at the x quadrature points (derivative_container.h:158-251)
void vp_set_field(hdo_problem &P, const XSpace &X, const double *phi)
{
  const int dx = X.dx, nd = X.nd, nq = X.nq;
  const int nqx = ipow(nq, dx);
  P.efield.assign((size_t)X.n_cells * nqx * dx, 0.0);
  std::vector<double> g(nqx);
  for (long long c = 0; c < X.n_cells; ++c)
    for (int e = 0; e < dx; ++e)
      {
        std::vector<const double *> A(dx);
        std::vector<int> m(dx, nq);
        for (int j = 0; j < dx; ++j)
          A[j] = j == e ? X.G.data() : X.S.data();
        tensor_apply(dx, nd, A, m, phi + c * X.dofs_cell, g.data());
        for (int q = 0; q < nqx; ++q)
          P.efield[((size_t)c * nqx + q) * dx + e] = g[q] / X.h[e];
      }
}

// sum_K sum_q (d_e phi)^2 JxW per component (diagnostics.h:95-141)
void vp_energy(const hdo_problem &P, const XSpace &X, double *out)
{
  const int dx = X.dx, nq = X.nq, nqx = ipow(nq, dx);
  for (int e = 0; e < dx; ++e)
    out[e] = 0.0;
  for (long long c = 0; c < X.n_cells; ++c)
    for (int q = 0; q < nqx; ++q)
      {
        double jxw = 1.0;
        int rem = q;
        for (int j = 0; j < dx; ++j)
          {
            jxw *= X.h[j] * X.w[rem % nq];
            rem /= nq;
          }
        for (int e = 0; e < dx; ++e)
          {
            const double g = P.efield[((size_t)c * nqx + q) * dx + e];
            out[e] += g * g * jxw;
          }
      }
}
} // namespace

// ---------------------------------------------------------------------------
// C interface
// ---------------------------------------------------------------------------
extern "C" {

hdo_problem *hdo_create(int dim_x, int dim_v, int degree, int n_points,
                        int collocation, const int *n_cells, const double *lo,
                        const double *hi, int periodic_x, int periodic_v,
                        double skew, const double *a, int bc_kind, int func_id,
                        const double *func_params)
{
  if (dim_x < 1 || dim_x > 3 || dim_v < 1 || dim_v > 3 || degree < 1 ||
      n_points < degree + 1 || (collocation && n_points != degree + 1))
    return nullptr;
  auto *P = new hdo_problem();
  P->dim_x = dim_x;
  P->dim_v = dim_v;
  P->D = dim_x + dim_v;
  P->degree = degree;
  P->nd = degree + 1;
  P->nq = n_points;
  P->collocation = collocation != 0;
  P->skew = skew;
  P->bc_kind = bc_kind;
  P->func_id = func_id;
  P->n_cells = 1;
  for (int d = 0; d < P->D; ++d)
    {
      P->n[d] = n_cells[d];
      P->lo[d] = lo[d];
      P->hi[d] = hi[d];
      P->h[d] = (hi[d] - lo[d]) / n_cells[d];
      P->periodic[d] = d < dim_x ? periodic_x != 0 : periodic_v != 0;
      P->a[d] = a ? a[d] : 0.0;
      P->n_cells *= n_cells[d];
    }
  P->func_params.assign(func_params ? func_params : P->a,
                        (func_params ? func_params : P->a) + P->D);
  if (func_id == HDO_FUNC_LANDAU && func_params)
    P->func_params.assign(func_params, func_params + 3);
  P->dofs_cell = ipow(P->nd, P->D);
  P->t.build(degree, n_points, P->collocation);
  P->tg.build(degree, degree + 1, false);

  const int D = P->D, nq = P->nq;
  P->jxw.assign(ipow(nq, D), 0.0);
  for (int q = 0; q < ipow(nq, D); ++q)
    {
      double v = 1.0;
      int rem = q;
      for (int j = 0; j < D; ++j)
        {
          v *= P->h[j] * P->t.wq[rem % nq];
          rem /= nq;
        }
      P->jxw[q] = v;
    }
  for (int d = 0; d < D; ++d)
    {
      P->jxwf[d].assign(ipow(nq, D - 1), 0.0);
      for (int q = 0; q < ipow(nq, D - 1); ++q)
        {
          double v = 1.0;
          int rem = q;
          for (int j = 0; j < D; ++j)
            {
              if (j == d)
                continue;
              v *= P->h[j] * P->t.wq[rem % nq];
              rem /= nq;
            }
          P->jxwf[d][q] = v;
        }
    }
  P->qidx.assign((size_t)ipow(nq, D) * D, 0);
  for (int q = 0; q < ipow(nq, D); ++q)
    {
      int rem = q;
      for (int j = 0; j < D; ++j)
        {
          P->qidx[(size_t)q * D + j] = rem % nq;
          rem /= nq;
        }
    }
  const Tables1D &tg = P->collocation ? P->t : P->tg;
  const int nd = P->nd;
  P->jxw_g.assign(ipow(nd, D), 0.0);
  for (int q = 0; q < ipow(nd, D); ++q)
    {
      double v = 1.0;
      int rem = q;
      for (int j = 0; j < D; ++j)
        {
          v *= P->h[j] * tg.wq[rem % nd];
          rem /= nd;
        }
      P->jxw_g[q] = v;
    }
  return P;
}

void hdo_destroy(hdo_problem *p) { delete p; }

long long hdo_n_dofs(const hdo_problem *p) { return p->n_cells * p->dofs_cell; }
long long hdo_n_cells(const hdo_problem *p) { return p->n_cells; }

int hdo_apply(hdo_problem *p, const double *src, double *dst, double t,
              int use_fcl)
{
  if (!p || !src || !dst)
    return -1;
  if (use_fcl)
    apply_fcl(*p, src, dst, t);
  else
    apply_ecl(*p, src, dst, t);
  return 0;
}

int hdo_apply_cells(hdo_problem *p, const double *src, double *dst,
                    double t, const long long *cell_ids, int n)
{
  if (!p || !src || !dst || !cell_ids || n < 0)
    return -1;
  for (int i = 0; i < n; ++i)
    if (cell_ids[i] < 0 || cell_ids[i] >= p->n_cells)
      return -1;
  const int nm = (n + L - 1) / L;
#pragma omp parallel
  {
    Work w(*p);
#pragma omp for schedule(static)
    for (int m = 0; m < nm; ++m)
      {
        long long cells[L];
        bool valid[L];
        for (int l = 0; l < L; ++l)
          {
            const int i = m * L + l;
            valid[l] = i < n;
            cells[l] = cell_ids[valid[l] ? i : n - 1];
          }
        ecl_cells(*p, w, src, dst, t, cells, valid);
      }
  }
  return 0;
}

int hdo_interpolate(hdo_problem *p, double t, double *dst)
{
  if (!p || !dst)
    return -1;
  // vector_tools.h:81-135: function values at the Gauss-Lobatto nodes
  const hdo_problem &P = *p;
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < P.n_cells; ++c)
    {
      int idx[6];
      P.cell_index(c, idx);
      for (long long i = 0; i < P.dofs_cell; ++i)
        {
          double xh[6], x[6];
          long long rem = i;
          for (int j = 0; j < P.D; ++j)
            {
              xh[j] = P.lo[j] + (idx[j] + P.t.xd[rem % P.nd]) * P.h[j];
              rem /= P.nd;
            }
          map_point(P, xh, x);
          dst[c * P.dofs_cell + i] =
            eval_function(P.func_id, P.func_params, P.dim_x, P.D, x, t);
        }
    }
  return 0;
}

int hdo_norm_and_error(hdo_problem *p, const double *src, double t,
                       double *out)
{
  if (!p || !src || !out)
    return -1;
  // vector_tools.h:144-223: interpolate to the quadrature points, then
  // sum f^2 JxW and (f - f_ex)^2 JxW
  const hdo_problem &P = *p;
  const int D = P.D, nd = P.nd, nq = P.nq;
  const long long nm = n_macro(P);
  double s0 = 0.0, s1 = 0.0;
#pragma omp parallel reduction(+ : s0, s1)
  {
    Work w(P);
#pragma omp for schedule(static)
    for (long long m = 0; m < nm; ++m)
      {
        long long cells[L];
        bool valid[L];
        int idx[L][6];
        macro_cells(P, m, cells, valid, idx);
        for (long long i = 0; i < P.dofs_cell; ++i)
          for (int l = 0; l < L; ++l)
            w.u[i * L + l] = src[cells[l] * P.dofs_cell + i];
        int ext[6];
        for (int j = 0; j < D; ++j)
          ext[j] = nd;
        double *uq = w.u.data();
        if (!P.collocation)
          uq = contract_dirs(P, ext, -1, nq, P.t.S.data(), nd, 1, w.u.data(),
                             w.uq.data(), false);
        const int nvol = ipow(nq, D);
        for (int q = 0; q < nvol; ++q)
          for (int l = 0; l < L; ++l)
            {
              if (!valid[l])
                continue;
              double xh[6], x[6];
              int rem = q;
              for (int j = 0; j < D; ++j)
                {
                  xh[j] = P.lo[j] + (idx[l][j] + P.t.xq[rem % nq]) * P.h[j];
                  rem /= nq;
                }
              map_point(P, xh, x);
              const double jxw =
                P.deformed
                  ? cell_jxw(P, idx[l], &P.qidx[(size_t)q * D], P.t.xq, P.t.wq)
                  : P.jxw[q];
              const double f = uq[q * L + l];
              const double e =
                f - eval_function(P.func_id, P.func_params, P.dim_x, D, x, t);
              s0 += f * f * jxw;
              s1 += e * e * jxw;
            }
      }
  }
  out[0] = std::sqrt(s0);
  out[1] = std::sqrt(s1);
  return 0;
}

double hdo_critical_time_step(const hdo_problem *p)
{
  // cfl.h:56-124: per space, the largest |(J^{-T} a)_j| over its cells and
  // Gauss points (n_points); Cartesian: a_d / h_d
  double vx = 0.0, vv = 0.0;
  if (!p->deformed)
    for (int d = 0; d < p->D; ++d)
      {
        const double v = std::fabs(p->a[d] / p->h[d]);
        if (d < p->dim_x)
          vx = std::max(vx, v);
        else
          vv = std::max(vv, v);
      }
  else
    {
      const hdo_problem &P = *p;
      std::vector<double> xg, wg;
      gauss_rule(P.nq, xg, wg);
      for (int s = 0; s < 2; ++s)
        {
          const int b = s == 0 ? 0 : P.dim_x, e = s == 0 ? P.dim_x : P.D;
          long long nc = 1;
          int nqs = 1;
          for (int j = b; j < e; ++j)
            {
              nc *= P.n[j];
              nqs *= P.nq;
            }
          double vmax = 0.0;
          for (long long c = 0; c < nc; ++c)
            for (int q = 0; q < nqs; ++q)
              {
                int idx[6] = {0, 0, 0, 0, 0, 0};
                double xi[6] = {0.5, 0.5, 0.5, 0.5, 0.5, 0.5};
                long long cr = c;
                int qr = q;
                for (int j = b; j < e; ++j)
                  {
                    idx[j] = (int)(cr % P.n[j]);
                    cr /= P.n[j];
                    xi[j] = xg[qr % P.nq];
                    qr /= P.nq;
                  }
                PointGeom g;
                point_geometry(P, idx, xi, g);
                for (int j = b; j < e; ++j)
                  {
                    double v = 0.0;
                    for (int i = b; i < e; ++i)
                      v += g.Ji[i][j] * P.a[i];
                    vmax = std::max(vmax, std::fabs(v));
                  }
              }
          (s == 0 ? vx : vv) = vmax;
        }
    }
  return std::min(1.0 / vx, 1.0 / vv);
}

int hdo_geometry(const hdo_problem *p, const int *cell_index,
                 const double *xi, double *x, double *inverse_jacobian,
                 double *det)
{
  if (!p || !cell_index || !xi)
    return -1;
  const hdo_problem &P = *p;
  if (x)
    {
      double xh[6];
      for (int j = 0; j < P.D; ++j)
        xh[j] = P.lo[j] + (cell_index[j] + xi[j]) * P.h[j];
      map_point(P, xh, x);
    }
  PointGeom g;
  point_geometry(P, cell_index, xi, g);
  if (inverse_jacobian)
    for (int i = 0; i < P.D; ++i)
      for (int j = 0; j < P.D; ++j)
        inverse_jacobian[i * P.D + j] = g.Ji[i][j];
  if (det)
    *det = g.det;
  return 0;
}

int hdo_set_deformation(hdo_problem *p, double deformation_x, int frequency_x,
                        double deformation_v, int frequency_v)
{
  if (!p || p->vp || frequency_x < 1 || frequency_v < 1)
    return -1;
  p->deform[0] = deformation_x;
  p->deform[1] = deformation_v;
  p->freq[0] = frequency_x;
  p->freq[1] = frequency_v;
  p->deformed = deformation_x != 0.0 || deformation_v != 0.0;
  return 0;
}

int hdo_lsrk_n_stages(const char *rk_type)
{
  std::vector<double> a, b;
  if (!rk_coefficients(rk_type, a, b))
    return -1;
  return (int)b.size();
}

int hdo_lsrk_coefficients(const char *rk_type, double *a, double *b)
{
  std::vector<double> va, vb;
  if (!rk_coefficients(rk_type, va, vb))
    return -1;
  std::copy(va.begin(), va.end(), a);
  std::copy(vb.begin(), vb.end(), b);
  return (int)vb.size();
}

int hdo_lsrk_step(hdo_problem *p, const char *rk_type, double *sol, double *ki,
                  double *ti, double t, double dt, int use_ecl)
{
  std::vector<double> a, b;
  if (!p || !rk_coefficients(rk_type, a, b))
    return -1;
  const size_t n = (size_t)hdo_n_dofs(p);
  Op op = [&](const double *src, double *dst, double time) {
    hdo_apply(p, src, dst, time, use_ecl ? 0 : 1);
  };
  lsrk_step(a, b, n, sol, ki, ti, t, dt, op);
  return 0;
}

double hdo_lsrk_ode(const char *rk_type, double dt, int n_steps, double y0)
{
  std::vector<double> a, b;
  if (!rk_coefficients(rk_type, a, b))
    return NAN;
  double y = y0, k = 0.0, ti = 0.0;
  Op op = [](const double *src, double *dst, double time) {
    const double s = std::sin(time);
    dst[0] = src[0] * s * s;
  };
  for (int it = 0; it < n_steps; ++it)
    lsrk_step(a, b, 1, &y, &k, &ti, dt * it, dt, op);
  return y;
}

int hdo_run_advection(hdo_problem *p, const char *rk_type, double time_step,
                      double start_time, double final_time, double cfl,
                      double diag_tick, int use_ecl, int max_steps,
                      double *rows, int max_rows, double *dt_used)
{
  std::vector<double> a, b;
  if (!p || !rk_coefficients(rk_type, a, b))
    return -1;
  const size_t n = (size_t)hdo_n_dofs(p);
  std::vector<double> sol(n), ki(n), ti(n);
  hdo_interpolate(p, start_time, sol.data());

  // application.h:363-396
  const double crit = hdo_critical_time_step(p);
  double dt = std::min(time_step, cfl * crit / std::pow((double)p->degree, 1.5));
  dt = (final_time - start_time) / std::ceil((final_time - start_time) / dt);
  if (dt_used)
    *dt_used = dt;

  int n_rows = 0;
  auto diagnostics = [&](double cur_time) {
    // application.h:467-476: print at the start and whenever the time
    // crosses a diagnostics tick
    if (cur_time != start_time &&
        (int)((cur_time + 0.00000000001 - start_time) / diag_tick) ==
          (int)((cur_time + 0.00000000001 - start_time - dt) / diag_tick))
      return;
    double e[2];
    hdo_norm_and_error(p, sol.data(), cur_time, e);
    if (n_rows < max_rows)
      {
        rows[3 * n_rows + 0] = cur_time;
        rows[3 * n_rows + 1] = e[0];
        rows[3 * n_rows + 2] = e[1];
      }
    ++n_rows;
  };
  Op op = [&](const double *src, double *dst, double time) {
    hdo_apply(p, src, dst, time, use_ecl ? 0 : 1);
  };

  // time_loop.cc:33-67
  diagnostics(start_time);
  int step = 1;
  for (double time = start_time + dt;
       time <= final_time * 1.0000000000001 && step <= max_steps;
       time += dt, ++step)
    {
      lsrk_step(a, b, n, sol.data(), ki.data(), ti.data(), time - dt, dt, op);
      diagnostics(time);
    }
  return std::min(n_rows, max_rows);
}

int hdo_tables(int degree, int n_points, int collocation, double *nodes,
               double *qpoints, double *qweights, double *S, double *D,
               double *F, double *Sinv)
{
  Tables1D t;
  t.build(degree, n_points, collocation != 0);
  std::copy(t.xd.begin(), t.xd.end(), nodes);
  std::copy(t.xq.begin(), t.xq.end(), qpoints);
  std::copy(t.wq.begin(), t.wq.end(), qweights);
  std::copy(t.S.begin(), t.S.end(), S);
  std::copy(t.D.begin(), t.D.end(), D);
  std::copy(t.F[0].begin(), t.F[0].end(), F);
  std::copy(t.F[1].begin(), t.F[1].end(), F + t.nq);
  std::copy(t.Sinv.begin(), t.Sinv.end(), Sinv);
  return 0;
}

double hdo_splitmix_uniform(unsigned long long seed, unsigned long long i)
{
  // SplitMix64 (Steele, Lea, Flood 2014) output for the state after i+1
  // increments of the golden-ratio constant, mapped to [-1, 1)
  uint64_t z = seed + (i + 1ULL) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

int hdo_add_noise(const hdo_problem *p, double *dst, unsigned long long seed,
                  double amplitude)
{
  const long long n = hdo_n_dofs(p);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < n; ++i)
    dst[i] += amplitude * hdo_splitmix_uniform(seed, (unsigned long long)i);
  return 0;
}

int hdo_set_vlasov_poisson(hdo_problem *p, int enable)
{
  if (p && p->deformed && enable)
    return -1;
  if (!p || (enable && p->dim_x != p->dim_v))
    return -1; // the reference's v-integration assumes dim_x == dim_v
  p->vp = enable != 0;
  if (p->vp)
    {
      XSpace X = make_xspace(*p);
      p->efield.assign((size_t)X.n_cells * ipow(p->nq, p->dim_x) * p->dim_x,
                       0.0);
    }
  return 0;
}

long long hdo_x_n_dofs(const hdo_problem *p)
{
  long long n = 1;
  for (int d = 0; d < p->dim_x; ++d)
    n *= p->n[d];
  return n * ipow(p->nd, p->dim_x);
}

int hdo_vp_density(hdo_problem *p, const double *f, double *rho)
{
  if (!p || !f || !rho)
    return -1;
  vp_density(*p, make_xspace(*p), f, rho);
  return 0;
}

int hdo_vp_rhs(hdo_problem *p, double *rho, double *b)
{
  if (!p || !rho || !b)
    return -1;
  vp_rhs(make_xspace(*p), rho, b);
  return 0;
}

int hdo_poisson_apply(hdo_problem *p, const double *src, double *dst)
{
  if (!p || !src || !dst)
    return -1;
  sipg_apply(make_xspace(*p), p->degree, src, dst);
  return 0;
}

int hdo_poisson_solve(hdo_problem *p, const double *b, double *phi,
                      double rel_tol, int max_it)
{
  if (!p || !b || !phi)
    return -1;
  return poisson_cg(make_xspace(*p), p->degree, b, phi, rel_tol, max_it);
}

int hdo_vp_set_potential(hdo_problem *p, const double *phi)
{
  if (!p || !phi || !p->vp)
    return -1;
  vp_set_field(*p, make_xspace(*p), phi);
  return 0;
}

int hdo_vp_efield(const hdo_problem *p, double *out)
{
  if (!p || !out || !p->vp)
    return -1;
  std::copy(p->efield.begin(), p->efield.end(), out);
  return 0;
}

int hdo_vp_energy(hdo_problem *p, double *out)
{
  if (!p || !out || !p->vp)
    return -1;
  vp_energy(*p, make_xspace(*p), out);
  return 0;
}

int hdo_vp_diagnostics(hdo_problem *p, const double *f, double *out)
{
  // mass, l2 norm, kinetic energy, momentum (3) at the quadrature points
  // (examples/vlasov_poisson/include/diagnostics.h:28-87)
  if (!p || !f || !out)
    return -1;
  const hdo_problem &P = *p;
  const int D = P.D, nd = P.nd, nq = P.nq, dv = P.dim_v;
  const long long nm = n_macro(P);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (long long m = 0; m < nm; ++m)
    {
      Work w(P);
      long long cells[L];
      bool valid[L];
      int idx[L][6];
      macro_cells(P, m, cells, valid, idx);
      for (long long i = 0; i < P.dofs_cell; ++i)
        for (int l = 0; l < L; ++l)
          w.u[i * L + l] = f[cells[l] * P.dofs_cell + i];
      int ext[6];
      for (int j = 0; j < D; ++j)
        ext[j] = nd;
      double *uq = w.u.data();
      if (!P.collocation)
        uq = contract_dirs(P, ext, -1, nq, P.t.S.data(), nd, 1, w.u.data(),
                           w.uq.data(), false);
      const int nvol = ipow(nq, D);
      for (int q = 0; q < nvol; ++q)
        for (int l = 0; l < L; ++l)
          {
            if (!valid[l])
              continue;
            const double fv = uq[q * L + l], jxw = P.jxw[q];
            double v2 = 0.0, vd[3] = {0, 0, 0};
            for (int e = 0; e < dv; ++e)
              {
                const int j = P.dim_x + e;
                vd[e] = P.lo[j] +
                        (idx[l][j] + P.t.xq[P.qidx[(size_t)q * D + j]]) *
                          P.h[j];
                v2 += vd[e] * vd[e];
              }
            acc[0] += fv * jxw;
            acc[1] += fv * fv * jxw;
            acc[2] += v2 * fv * jxw;
            for (int e = 0; e < dv; ++e)
              acc[3 + e] += vd[e] * fv * jxw;
          }
    }
  acc[1] = std::sqrt(acc[1]);
  std::copy(acc, acc + 6, out);
  return 0;
}

int hdo_run_vlasov_poisson(hdo_problem *p, const char *rk_type,
                           double time_step, double start_time,
                           double final_time, double cfl, double diag_tick,
                           double poisson_tol, int max_steps, double *rows,
                           int max_rows, double *dt_used, int *cg_iterations)
{
  // vp/application.h:470-660: interpolate f0, dt from the CFL rule with the
  // case's transport direction (p->a), LSRK with the RHS rho -> b -> CG ->
  // grad phi -> advection, diagnostics rows (t, energy (dim_x), mass, l2,
  // kinetic, momentum (3))
  std::vector<double> a, b;
  if (!p || !rk_coefficients(rk_type, a, b) || hdo_set_vlasov_poisson(p, 1))
    return -1;
  XSpace X = make_xspace(*p);
  const size_t n = (size_t)hdo_n_dofs(p);
  std::vector<double> sol(n), ki(n), ti(n);
  std::vector<double> rho(X.n_dofs), rhs(X.n_dofs), phi(X.n_dofs, 0.0);
  hdo_interpolate(p, start_time, sol.data());
  const double crit = hdo_critical_time_step(p);
  double dt = std::min(time_step, cfl * crit / std::pow((double)p->degree, 1.5));
  dt = (final_time - start_time) / std::ceil((final_time - start_time) / dt);
  if (dt_used)
    *dt_used = dt;
  long long total_it = 0;
  const int width = 1 + p->dim_x + 6;
  int n_rows = 0;
  auto diagnostics = [&](double cur_time) {
    if (cur_time != start_time &&
        (int)((cur_time + 0.00000000001 - start_time) / diag_tick) ==
          (int)((cur_time + 0.00000000001 - start_time - dt) / diag_tick))
      return;
    if (n_rows < max_rows)
      {
        double *row = rows + (size_t)n_rows * width;
        row[0] = cur_time;
        vp_energy(*p, X, row + 1); // field of the last RHS evaluation
        hdo_vp_diagnostics(p, sol.data(), row + 1 + p->dim_x);
      }
    ++n_rows;
  };
  Op op = [&](const double *src, double *dst, double time) {
    vp_density(*p, X, src, rho.data());
    vp_rhs(X, rho.data(), rhs.data());
    total_it += poisson_cg(X, p->degree, rhs.data(), phi.data(), poisson_tol,
                           100000);
    vp_set_field(*p, X, phi.data());
    apply_ecl(*p, src, dst, time);
  };
  diagnostics(start_time);
  int step = 1;
  for (double time = start_time + dt;
       time <= final_time * 1.0000000000001 && step <= max_steps;
       time += dt, ++step)
    {
      lsrk_step(a, b, n, sol.data(), ki.data(), ti.data(), time - dt, dt, op);
      diagnostics(time);
    }
  if (cg_iterations)
    *cg_iterations = (int)std::min<long long>(total_it, 2147483647LL);
  return std::min(n_rows, max_rows);
}

int hdo_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

} // extern "C"
