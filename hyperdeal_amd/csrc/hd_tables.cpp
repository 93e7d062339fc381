// hd_tables.cpp -- 1D tables and the collapsed 1D operator (host code).
#include "hd_tables.h"

#include <cmath>

namespace hd
{
namespace
{
using ld = long double;
const ld kPi = 3.141592653589793238462643383279502884L;

// P_n(z) and P_n'(z) by the three-term recurrence
void legendre_pd(int n, ld z, ld &p, ld &dp)
{
  ld pm = 1.0L, pc = z;
  if (n == 0)
    {
      p = 1.0L;
      dp = 0.0L;
      return;
    }
  for (int j = 1; j < n; ++j)
    {
      const ld pn = ((2 * j + 1) * z * pc - j * pm) / (j + 1);
      pm = pc;
      pc = pn;
    }
  p = pc;
  dp = n * (pm - z * pc) / (1.0L - z * z);
}

// roots of P_n on [-1,1] (Newton from Chebyshev guesses), weights 2/((1-z^2)P'^2)
void gauss(int n, std::vector<double> &x, std::vector<double> &w)
{
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i)
    {
      ld z = -std::cos(kPi * (i + 0.75L) / (n + 0.5L));
      ld p = 0, dp = 1;
      for (int it = 0; it < 60; ++it)
        {
          legendre_pd(n, z, p, dp);
          const ld dz = p / dp;
          z -= dz;
          if (std::fabs((double)dz) < 1e-20)
            break;
        }
      legendre_pd(n, z, p, dp);
      x[i] = (double)((1.0L + z) / 2.0L);
      w[i] = (double)(1.0L / ((1.0L - z * z) * dp * dp));
    }
}

// endpoints and the roots of P_{n-1}'; weights 2/(n(n-1) P_{n-1}^2)
void gauss_lobatto(int n, std::vector<double> &x, std::vector<double> &w)
{
  x.resize(n);
  w.resize(n);
  const int m = n - 1;
  for (int i = 0; i < n; ++i)
    {
      ld z;
      if (i == 0)
        z = -1.0L;
      else if (i == m)
        z = 1.0L;
      else
        {
          z = -std::cos(kPi * i / m);
          for (int it = 0; it < 60; ++it)
            {
              ld p, dp;
              legendre_pd(m, z, p, dp);
              const ld d2p = (2.0L * z * dp - m * (m + 1.0L) * p) /
                             (1.0L - z * z);
              const ld dz = dp / d2p;
              z -= dz;
              if (std::fabs((double)dz) < 1e-20)
                break;
            }
        }
      ld p, dp;
      if (i == 0 || i == m)
        p = (i == 0 && (m % 2 == 1)) ? -1.0L : 1.0L;
      else
        legendre_pd(m, z, p, dp);
      x[i] = (double)((1.0L + z) / 2.0L);
      w[i] = (double)(1.0L / (n * (n - 1.0L) * p * p));
    }
}

// barycentric weights of a node set
std::vector<ld> bary_weights(const std::vector<double> &xs)
{
  std::vector<ld> bw(xs.size(), 1.0L);
  for (size_t j = 0; j < xs.size(); ++j)
    for (size_t m = 0; m < xs.size(); ++m)
      if (m != j)
        bw[j] /= ((ld)xs[j] - xs[m]);
  return bw;
}

// value of the j-th Lagrange polynomial on xs at x
ld lagrange(const std::vector<double> &xs, const std::vector<ld> &bw, int j,
            ld x)
{
  ld prod = 1.0L;
  for (size_t m = 0; m < xs.size(); ++m)
    {
      if ((int)m == j)
        continue;
      prod *= (x - xs[m]);
    }
  return prod * bw[j];
}

// differentiation matrix of the Lagrange basis on xs at its own points:
// D[q][p] = l_p'(x_q)  (off-diagonal w_p/w_q/(x_q-x_p), diagonal = -row sum)
std::vector<double> diff_matrix(const std::vector<double> &xs)
{
  const int n = (int)xs.size();
  const auto bw = bary_weights(xs);
  std::vector<double> D(n * n, 0.0);
  for (int q = 0; q < n; ++q)
    {
      ld diag = 0.0L;
      for (int p = 0; p < n; ++p)
        {
          if (p == q)
            continue;
          const ld v = (bw[p] / bw[q]) / ((ld)xs[q] - xs[p]);
          D[q * n + p] = (double)v;
          diag -= v;
        }
      D[q * n + q] = (double)diag;
    }
  return D;
}

// Gauss-Jordan solve of the small SPD system M X = R (row-major)
void solve(int n, int m, std::vector<ld> &M, std::vector<ld> &R)
{
  for (int c = 0; c < n; ++c)
    {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (std::fabs((double)M[r * n + c]) > std::fabs((double)M[piv * n + c]))
          piv = r;
      for (int j = 0; j < n && piv != c; ++j)
        std::swap(M[c * n + j], M[piv * n + j]);
      for (int j = 0; j < m && piv != c; ++j)
        std::swap(R[c * m + j], R[piv * m + j]);
      const ld inv = 1.0L / M[c * n + c];
      for (int j = 0; j < n; ++j)
        M[c * n + j] *= inv;
      for (int j = 0; j < m; ++j)
        R[c * m + j] *= inv;
      for (int r = 0; r < n; ++r)
        {
          if (r == c)
            continue;
          const ld f = M[r * n + c];
          for (int j = 0; j < n; ++j)
            M[r * n + j] -= f * M[c * n + j];
          for (int j = 0; j < m; ++j)
            R[r * m + j] -= f * R[c * m + j];
        }
    }
}
} // namespace

bool Tables1D::build(int degree, int n_points, bool coll)
{
  if (degree < 1 || degree > 15 || n_points < degree + 1 || n_points > 16 ||
      (coll && n_points != degree + 1))
    return false;
  nd = degree + 1;
  nq = n_points;
  collocation = coll;
  gauss_lobatto(nd, nodes, node_weights);
  if (coll)
    gauss_lobatto(nq, qpoints, qweights);
  else
    gauss(nq, qpoints, qweights);

  const auto bwn = bary_weights(nodes);
  const auto bwq = bary_weights(qpoints);
  S.assign(nq * nd, 0.0);
  for (int q = 0; q < nq; ++q)
    for (int i = 0; i < nd; ++i)
      S[q * nd + i] = coll ? (q == i ? 1.0 : 0.0)
                           : (double)lagrange(nodes, bwn, i, qpoints[q]);
  Dq = diff_matrix(qpoints);
  // derivatives of the nodal basis: at the quadrature points and the ends
  {
    const auto dnodal = [&](int i, ld x) {
      ld sum = 0.0L;
      for (int m = 0; m < nd; ++m)
        {
          if (m == i)
            continue;
          ld prod = 1.0L / ((ld)nodes[i] - nodes[m]);
          for (int l = 0; l < nd; ++l)
            if (l != i && l != m)
              prod *= (x - nodes[l]) / ((ld)nodes[i] - nodes[l]);
          sum += prod;
        }
      return sum;
    };
    G.assign(nq * nd, 0.0);
    for (int q = 0; q < nq; ++q)
      for (int i = 0; i < nd; ++i)
        G[q * nd + i] = (double)dnodal(i, qpoints[q]);
    dnode0.assign(nd, 0.0);
    dnode1.assign(nd, 0.0);
    for (int i = 0; i < nd; ++i)
      {
        dnode0[i] = (double)dnodal(i, 0.0L);
        dnode1[i] = (double)dnodal(i, 1.0L);
      }
  }
  F.assign(2 * nq, 0.0);
  for (int s = 0; s < 2; ++s)
    for (int p = 0; p < nq; ++p)
      F[s * nq + p] = coll ? ((s == 0 ? p == 0 : p == nq - 1) ? 1.0 : 0.0)
                           : (double)lagrange(qpoints, bwq, p, (ld)s);
  // Sinv = (S^T W S)^{-1} S^T W
  std::vector<ld> M(nd * nd, 0.0L), R(nd * nq, 0.0L);
  for (int i = 0; i < nd; ++i)
    for (int q = 0; q < nq; ++q)
      {
        R[i * nq + q] = (ld)S[q * nd + i] * qweights[q];
        for (int j = 0; j < nd; ++j)
          M[i * nd + j] += (ld)S[q * nd + i] * qweights[q] * S[q * nd + j];
      }
  solve(nd, nq, M, R);
  Sinv.assign(nd * nq, 0.0);
  for (int i = 0; i < nd * nq; ++i)
    Sinv[i] = coll ? ((i / nq) == (i % nq) ? 1.0 : 0.0) : (double)R[i];
  return true;
}

namespace
{
// The 1D ECL algorithm on one cell [0, h] (long double), u nodal values,
// ul / ur the left / right neighbours' face values (their right / left node).
void apply_1d(const Tables1D &t, ld a, ld h, ld skew, const ld *u, ld ul,
              ld ur, ld *out)
{
  const int nd = t.nd, nq = t.nq;
  std::vector<ld> uq(nq, 0.0L), data(nq, 0.0L), tmp(nq, 0.0L);
  for (int q = 0; q < nq; ++q)
    for (int i = 0; i < nd; ++i)
      uq[q] += (ld)t.S[q * nd + i] * u[i];
  // cell: -skew (a/h)(D uq) h w + D^T[(1-skew) uq (a/h) h w]
  for (int q = 0; q < nq; ++q)
    {
      ld g = 0.0L;
      for (int p = 0; p < nq; ++p)
        g += (ld)t.Dq[q * nq + p] * uq[p];
      data[q] = -skew * (a / h) * g * h * t.qweights[q];
      tmp[q] = (1.0L - skew) * uq[q] * (a / h) * h * t.qweights[q];
    }
  for (int p = 0; p < nq; ++p)
    for (int q = 0; q < nq; ++q)
      data[p] += (ld)t.Dq[q * nq + p] * tmp[q];
  // faces: side 0 (normal -1, neighbour ul), side 1 (normal +1, ur)
  for (int s = 0; s < 2; ++s)
    {
      ld um = 0.0L;
      for (int p = 0; p < nq; ++p)
        um += (ld)t.F[s * nq + p] * uq[p];
      const ld up = s ? ur : ul;
      const ld an = a * (s ? 1.0L : -1.0L);
      const ld flux = 0.5L * ((um + up) * an + std::fabs(an) * (um - up));
      const ld fd = -(flux - skew * um * an); // face JxW of a point is 1
      for (int p = 0; p < nq; ++p)
        data[p] += (ld)t.F[s * nq + p] * fd;
    }
  for (int q = 0; q < nq; ++q)
    data[q] /= h * t.qweights[q];
  for (int i = 0; i < nd; ++i)
    {
      out[i] = 0.0L;
      for (int q = 0; q < nq; ++q)
        out[i] += (ld)t.Sinv[i * nq + q] * data[q];
    }
}
} // namespace

Collapsed1D collapse_direction(const Tables1D &t, double a, double h,
                               double skew)
{
  const int nd = t.nd;
  Collapsed1D c;
  c.A.assign(nd * nd, 0.0);
  c.b.assign(nd, 0.0);
  c.inflow_from_left = a >= 0.0;
  std::vector<ld> u(nd, 0.0L), out(nd, 0.0L);
  for (int j = 0; j < nd; ++j)
    {
      std::fill(u.begin(), u.end(), 0.0L);
      u[j] = 1.0L;
      apply_1d(t, a, h, skew, u.data(), 0.0L, 0.0L, out.data());
      for (int i = 0; i < nd; ++i)
        c.A[i * nd + j] = (double)out[i];
    }
  std::fill(u.begin(), u.end(), 0.0L);
  apply_1d(t, a, h, skew, u.data(), c.inflow_from_left ? 1.0L : 0.0L,
           c.inflow_from_left ? 0.0L : 1.0L, out.data());
  for (int i = 0; i < nd; ++i)
    c.b[i] = (double)out[i];
  return c;
}
} // namespace hd
