// hd_tables.h -- host-side 1D operator tables for the phase-space DG kernels.
//
// hyper.deal takes these from deal.II's ShapeInfo (shape_values_eo,
// shape_gradients_collocation_eo, inverse_shape_values_eo,
// quadrature_data_on_face; used at operators/advection/advection_operation.h:
// 241-282).  They are rebuilt here from the mathematics: Lagrange bases on
// the k+1 Gauss-Lobatto nodes (the FE_DGQ support points) and on the
// quadrature points (Gauss, or Gauss-Lobatto for collocation).
#pragma once

#include <vector>

namespace hd
{
struct Tables1D
{
  int nd = 0;               // k+1 nodal points
  int nq = 0;               // quadrature points
  bool collocation = false; // quadrature points == nodes
  std::vector<double> nodes, qpoints, qweights;
  std::vector<double> S;    // nq x nd   nodal basis at quadrature points
  std::vector<double> Dq;   // nq x nq   derivative of the quadrature-point
                            //           Lagrange basis at quadrature points
  std::vector<double> F;    // 2 x nq    quadrature-point basis at 0 and 1
  std::vector<double> Sinv; // nd x nq   (S^T W S)^{-1} S^T W
  std::vector<double> node_weights; // Gauss-Lobatto weights of the nodes
  std::vector<double> dnode0, dnode1; // nodal basis derivatives at 0 and 1
  std::vector<double> G;    // nq x nd   nodal basis derivatives at the
                            //           quadrature points

  // returns false for unsupported (degree, n_points, collocation)
  bool build(int degree, int n_points, bool coll);
};

// The exact 1D factor of the constant-velocity Cartesian operator in one
// direction: for a cell of width h moving with velocity a,
//   out = A u + b u_nb
// where u are the cell's nodal values and u_nb the inflow neighbour's nodal
// value on the shared face (its last node for a >= 0, its first for a < 0).
// Built by applying the 1D version of the reference's ECL quadrature
// algorithm (advection_operation.h:221-566) to unit vectors, so it carries
// the same quadrature, skew form and inverse mass matrix.
struct Collapsed1D
{
  std::vector<double> A; // nd x nd, row-major (out index, in index)
  std::vector<double> b; // nd
  bool inflow_from_left = true;
};

Collapsed1D collapse_direction(const Tables1D &t, double a, double h,
                               double skew);
} // namespace hd
