// hd_geometry.h -- host-side geometry of a deformed phase-space mesh.
//
// hyper.deal deforms each space's Cartesian mesh with DeformedCubeManifold
// (source/grid/grid_generator.cc:110-230): a point x^ of the x (or v) space
// moves to x^ + delta prod_j sin(f pi (x^_j - lo_j) / (hi_j - lo_j)) (1..1).
// Here the map is used exactly (not through a polynomial mapping), and the
// phase-space Jacobian is block diagonal over the two spaces, so every
// per-point quantity of the operator is a product of an x-space and a
// v-space factor: the tables below are per space, per cell and per point
// (small: they scale with the number of x cells or v cells, not with their
// product).
#pragma once

#include <vector>

namespace hd
{
// one space (x or v) of the mesh
struct SpaceMap
{
  int dim = 0;
  int n[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, h[3] = {1, 1, 1};
  double delta = 0.0;
  int freq = 2;
  long long n_cells() const
  {
    long long c = 1;
    for (int d = 0; d < dim; ++d)
      c *= n[d];
    return c;
  }
};

// physical point of the reference-space point xh (in place)
void space_map(const SpaceMap &m, double *x);

// the cell map xi -> x of the cell with multi-index idx at reference point
// xi in [0,1]^dim: inverse Jacobian Ji[i][j] = d xi_i / d x_j and
// det(d x / d xi)
void space_jacobian(const SpaceMap &m, const int *idx, const double *xi,
                    double Ji[3][3], double &det);

// per-space tables of an operator with constant velocity a (dim values):
//   cell: [cell][point][dim + 1] = (J^{-1} a at the point, |det J| w)
//   face: [cell][2 dim][face point] = +-(J^{-1} a)_d |det J| w_face for the
//         face of direction d and side s (sign + for s = 1), i.e. (a . n)
//         times the face measure
// at the nq quadrature points (xq, wq) per direction; with a == nullptr only
// |det J| w (cell: [cell][point])
void space_tables(const SpaceMap &m, int nq, const double *xq,
                  const double *wq, const double *a, std::vector<double> &cell,
                  std::vector<double> &face);

// largest |(J^{-T} a)_j| over the cells and the points xq of the space
// (the quantity of the reference's compute_critical_time_step,
// operators/advection/cfl.h:56-90)
double space_max_speed(const SpaceMap &m, int nq, const double *xq,
                       const double *a);
} // namespace hd
