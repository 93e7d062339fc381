// hd_geometry.cpp -- host-side geometry of a deformed phase-space mesh (see
// hd_geometry.h).
#include "hd_geometry.h"

#include <cmath>

namespace hd
{
namespace
{
constexpr double kPi = 3.14159265358979323846;

// closed-form inverse of a 1x1, 2x2 or 3x3 matrix; returns the determinant
double invert_small(int dim, const double J[3][3], double Ji[3][3])
{
  if (dim == 1)
    {
      Ji[0][0] = 1.0 / J[0][0];
      return J[0][0];
    }
  if (dim == 2)
    {
      const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      Ji[0][0] = J[1][1] / det;
      Ji[0][1] = -J[0][1] / det;
      Ji[1][0] = -J[1][0] / det;
      Ji[1][1] = J[0][0] / det;
      return det;
    }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  Ji[0][0] = c00 / det;
  Ji[1][0] = c01 / det;
  Ji[2][0] = c02 / det;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  return det;
}

void cell_index(const SpaceMap &m, long long c, int *idx)
{
  for (int d = 0; d < m.dim; ++d)
    {
      idx[d] = (int)(c % m.n[d]);
      c /= m.n[d];
    }
}
} // namespace

void space_map(const SpaceMap &m, double *x)
{
  double s = m.delta;
  for (int j = 0; j < m.dim; ++j)
    s *= std::sin(m.freq * kPi * (x[j] - m.lo[j]) / (m.hi[j] - m.lo[j]));
  for (int j = 0; j < m.dim; ++j)
    x[j] += s;
}

void space_jacobian(const SpaceMap &m, const int *idx, const double *xi,
                    double Ji[3][3], double &det)
{
  double sn[3], cs[3];
  for (int j = 0; j < m.dim; ++j)
    {
      const double xh = m.lo[j] + (idx[j] + xi[j]) * m.h[j];
      const double arg = m.freq * kPi * (xh - m.lo[j]) / (m.hi[j] - m.lo[j]);
      sn[j] = std::sin(arg);
      cs[j] = std::cos(arg);
    }
  // x_i = x^_i + s(x^) for every i: d x_i / d x^_j = delta_ij + ds/dx^_j,
  // and d x^_j / d xi_j = h_j
  double J[3][3] = {};
  for (int j = 0; j < m.dim; ++j)
    {
      double ds = m.delta * m.freq * kPi / (m.hi[j] - m.lo[j]) * cs[j];
      for (int k = 0; k < m.dim; ++k)
        if (k != j)
          ds *= sn[k];
      for (int i = 0; i < m.dim; ++i)
        J[i][j] = ((i == j ? 1.0 : 0.0) + ds) * m.h[j];
    }
  det = invert_small(m.dim, J, Ji);
}

void space_tables(const SpaceMap &m, int nq, const double *xq,
                  const double *wq, const double *a, std::vector<double> &cell,
                  std::vector<double> &face)
{
  const int dim = m.dim;
  int npt = 1, nfpt = 1;
  for (int d = 0; d < dim; ++d)
    npt *= nq;
  for (int d = 1; d < dim; ++d)
    nfpt *= nq;
  const long long nc = m.n_cells();
  const int per = a ? dim + 1 : 1;
  cell.assign((size_t)nc * npt * per, 0.0);
  face.assign(a ? (size_t)nc * 2 * dim * nfpt : 0, 0.0);
  for (long long c = 0; c < nc; ++c)
    {
      int idx[3] = {0, 0, 0};
      cell_index(m, c, idx);
      for (int q = 0; q < npt; ++q)
        {
          double xi[3] = {0, 0, 0}, w = 1.0;
          int r = q;
          for (int j = 0; j < dim; ++j)
            {
              xi[j] = xq[r % nq];
              w *= wq[r % nq];
              r /= nq;
            }
          double Ji[3][3], det;
          space_jacobian(m, idx, xi, Ji, det);
          double *out = &cell[((size_t)c * npt + q) * per];
          if (a)
            for (int d = 0; d < dim; ++d)
              {
                double v = 0.0;
                for (int i = 0; i < dim; ++i)
                  v += Ji[d][i] * a[i];
                out[d] = v;
              }
          out[per - 1] = std::fabs(det) * w;
        }
      if (!a)
        continue;
      for (int d = 0; d < dim; ++d)
        for (int s = 0; s < 2; ++s)
          for (int p = 0; p < nfpt; ++p)
            {
              double xi[3] = {0, 0, 0}, w = 1.0;
              int r = p;
              for (int j = 0; j < dim; ++j)
                {
                  if (j == d)
                    {
                      xi[j] = s;
                      continue;
                    }
                  xi[j] = xq[r % nq];
                  w *= wq[r % nq];
                  r /= nq;
                }
              double Ji[3][3], det;
              space_jacobian(m, idx, xi, Ji, det);
              double v = 0.0;
              for (int i = 0; i < dim; ++i)
                v += Ji[d][i] * a[i];
              face[(((size_t)c * 2 * dim) + 2 * d + s) * nfpt + p] =
                (s ? 1.0 : -1.0) * v * std::fabs(det) * w;
            }
    }
}

double space_max_speed(const SpaceMap &m, int nq, const double *xq,
                       const double *a)
{
  int npt = 1;
  for (int d = 0; d < m.dim; ++d)
    npt *= nq;
  double vmax = 0.0;
  const long long nc = m.n_cells();
  for (long long c = 0; c < nc; ++c)
    {
      int idx[3] = {0, 0, 0};
      cell_index(m, c, idx);
      for (int q = 0; q < npt; ++q)
        {
          double xi[3] = {0, 0, 0};
          int r = q;
          for (int j = 0; j < m.dim; ++j)
            {
              xi[j] = xq[r % nq];
              r /= nq;
            }
          double Ji[3][3], det;
          space_jacobian(m, idx, xi, Ji, det);
          for (int j = 0; j < m.dim; ++j)
            {
              double v = 0.0;
              for (int i = 0; i < m.dim; ++i)
                v += Ji[i][j] * a[i];
              vmax = std::fmax(vmax, std::fabs(v));
            }
        }
    }
  return vmax;
}
} // namespace hd
