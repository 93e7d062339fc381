// The advection operator written against the host loop / evaluator
// interface (include/hyperdeal_amd_host.hpp) the way hyper.deal writes it:
// face-centric (cell, face and boundary operations for MatrixFree::loop and
// the inverse mass matrix for cell_loop; include/hyper.deal/operators/
// advection/advection_operation.h:573-1131) and element-centric (every face
// inside the cell operation, for loop_cell_centric; advection_operation.h:
// 221-566).  Constant velocity a, skew factor, homogeneous Dirichlet data
// (u+ = -u-) on boundary faces.
#pragma once

#include "hyperdeal_amd_host.hpp"

#include <array>
#include <cmath>
#include <vector>

namespace advection_host
{
using namespace hyperdeal_amd::host;

template <int dim_x, int dim_v>
struct Advection
{
  static constexpr int dim = dim_x + dim_v;
  using MF = MatrixFree<dim_x, dim_v>;
  using ID = TensorID;
  using Vec = std::vector<double>;

  const MF &mf;
  std::array<double, dim> a;
  double skew;
  int n, nq;

  Advection(const MF &mf, const std::array<double, dim> &a, double skew)
    : mf(mf), a(a), skew(skew), n(mf.n_points()), nq((int)mf.dofs_per_cell())
  {
  }

  static int ipow(int b, int e)
  {
    int r = 1;
    while (e-- > 0)
      r *= b;
    return r;
  }

  // the cell integrals at the points, in data (the values at the points on
  // entry); tempp: dim x nq scratch
  void cell_terms(FEEvaluation<dim_x, dim_v> &phi, double *data,
                  std::vector<double> &buffer, std::vector<double> &tempp)
  {
    const Shape1D &s = mf.get_shape_info();
    const TensorProduct eval_(dim, n, nullptr, &s.gradients, nullptr);
    for (int i = 0; i < nq; ++i)
      buffer[i] = data[i];
    const int nqx = ipow(n, dim_x), nqv = ipow(n, dim_v);
    // x space
    if (skew != 0.0)
      for (int d = 0; d < dim_x; ++d)
        eval_.apply(&s.gradients, d, true, false, buffer.data(),
                    tempp.data() + (size_t)nq * d);
    for (int qv = 0, q = 0; qv < nqv; ++qv)
      for (int qx = 0; qx < nqx; ++qx, ++q)
        {
          if (skew != 0.0)
            {
              const auto g = phi.get_gradient_x(tempp.data(), q, qx, qv);
              double gv = 0.0;
              for (int d = 0; d < dim_x; ++d)
                gv += g[d] * a[d];
              phi.template submit_value<false>(data, -skew * gv, q, qx, qv);
            }
          if (skew != 1.0)
            {
              std::array<double, dim_x> gi{};
              for (int d = 0; d < dim_x; ++d)
                gi[d] = (1.0 - skew) * buffer[q] * a[d];
              phi.submit_gradient_x(tempp.data(), gi, q, qx, qv);
            }
        }
    if (skew != 1.0)
      for (int d = 0; d < dim_x; ++d)
        eval_.apply(&s.gradients, d, false, d > 0 || skew != 0.0,
                    tempp.data() + (size_t)nq * d, data);
    // v space
    if (skew != 0.0)
      for (int d = 0; d < dim_v; ++d)
        eval_.apply(&s.gradients, dim_x + d, true, false, buffer.data(),
                    tempp.data() + (size_t)nq * d);
    for (int qv = 0, q = 0; qv < nqv; ++qv)
      for (int qx = 0; qx < nqx; ++qx, ++q)
        {
          if (skew != 0.0)
            {
              const auto g = phi.get_gradient_v(tempp.data(), q, qx, qv);
              double gv = 0.0;
              for (int d = 0; d < dim_v; ++d)
                gv += g[d] * a[dim_x + d];
              phi.template submit_value<true>(data, -skew * gv, q, qx, qv);
            }
          if (skew != 1.0)
            {
              std::array<double, dim_v> gi{};
              for (int d = 0; d < dim_v; ++d)
                gi[d] = (1.0 - skew) * buffer[q] * a[dim_x + d];
              phi.submit_gradient_v(tempp.data(), gi, q, qx, qv);
            }
        }
    if (skew != 1.0)
      for (int d = 0; d < dim_v; ++d)
        eval_.apply(&s.gradients, dim_x + d, false, true,
                    tempp.data() + (size_t)nq * d, data);
  }

  void values_all(double *data, bool forward, int dims) const
  {
    const Shape1D &s = mf.get_shape_info();
    const TensorProduct eval(dims, n, &s.values, nullptr, nullptr);
    for (int k = 0; k < dims; ++k)
      eval.apply(&s.values, forward ? k : dims - 1 - k, forward, false, data,
                 data);
  }

  // flux on a face at its points: submits to the minus side (and the plus
  // side when plus is given); boundary: u+ = -u- (homogeneous Dirichlet)
  void face_flux(FEFaceEvaluation<dim_x, dim_v> &phi_m, double *d1,
                 FEFaceEvaluation<dim_x, dim_v> *phi_p, double *d2,
                 bool boundary)
  {
    const bool in_x = phi_m.get_type() == TensorID::SpaceType::X;
    const int nqx = (int)phi_m.face_points_x();
    const int nqv = ipow(n, dim - 1) / nqx;
    for (int qv = 0, q = 0; qv < nqv; ++qv)
      for (int qx = 0; qx < nqx; ++qx, ++q)
        {
          const double um = d1[q];
          const double up = boundary ? -um : d2[q];
          double ns = 0.0;
          if (in_x)
            {
              const auto nv = phi_m.get_normal_vector_x(qx);
              for (int d = 0; d < dim_x; ++d)
                ns += a[d] * nv[d];
            }
          else
            {
              const auto nv = phi_m.get_normal_vector_v(qv);
              for (int d = 0; d < dim_v; ++d)
                ns += a[dim_x + d] * nv[d];
            }
          const double flux =
            0.5 * ((um + up) * ns + std::fabs(ns) * (um - up));
          if (in_x)
            phi_m.template submit_value<TensorID::SpaceType::X>(
              d1, flux - skew * um * ns, q, qx, qv);
          else
            phi_m.template submit_value<TensorID::SpaceType::V>(
              d1, flux - skew * um * ns, q, qx, qv);
          if (phi_p)
            {
              if (in_x)
                phi_p->template submit_value<TensorID::SpaceType::X>(
                  d2, -flux + skew * up * ns, q, qx, qv);
              else
                phi_p->template submit_value<TensorID::SpaceType::V>(
                  d2, -flux + skew * up * ns, q, qx, qv);
            }
        }
  }

  void inverse_mass(FEEvaluationInverse<dim_x, dim_v> &phi, double *data)
  {
    const Shape1D &s = mf.get_shape_info();
    const TensorProduct eval_inv(dim, n, nullptr, nullptr, &s.inverse);
    for (int d = 0; d < dim; ++d)
      eval_inv.apply(&s.inverse, d, true, false, data, data);
    const int nqx = ipow(n, dim_x), nqv = ipow(n, dim_v);
    for (int qv = 0, q = 0; qv < nqv; ++qv)
      for (int qx = 0; qx < nqx; ++qx, ++q)
        phi.submit_inv(data, q, qx, qv);
    for (int d = dim - 1; d >= 0; --d)
      eval_inv.apply(&s.inverse, d, false, false, data, data);
  }

  // face-centric
  void local_apply_cell(const MF &, Vec &dst, const Vec &src, const ID cell)
  {
    FEEvaluation<dim_x, dim_v> phi(mf);
    std::vector<double> buffer(nq), tempp((size_t)nq * dim);
    phi.reinit(cell);
    phi.read_dof_values(src);
    double *data = phi.get_data_ptr();
    values_all(data, true, dim);
    cell_terms(phi, data, buffer, tempp);
    values_all(data, false, dim);
    phi.set_dof_values(dst);
  }
  void local_apply_face(const MF &, Vec &dst, const Vec &src, const ID face)
  {
    FEFaceEvaluation<dim_x, dim_v> phi_m(mf, true), phi_p(mf, false);
    phi_m.reinit(face);
    phi_p.reinit(face);
    phi_m.read_dof_values(src);
    phi_p.read_dof_values(src);
    values_all(phi_m.get_data_ptr(), true, dim - 1);
    values_all(phi_p.get_data_ptr(), true, dim - 1);
    face_flux(phi_m, phi_m.get_data_ptr(), &phi_p, phi_p.get_data_ptr(),
              false);
    values_all(phi_m.get_data_ptr(), false, dim - 1);
    values_all(phi_p.get_data_ptr(), false, dim - 1);
    phi_m.distribute_local_to_global(dst);
    phi_p.distribute_local_to_global(dst);
  }
  void local_apply_boundary(const MF &, Vec &dst, const Vec &src,
                            const ID face)
  {
    FEFaceEvaluation<dim_x, dim_v> phi_m(mf, true);
    phi_m.reinit(face);
    phi_m.read_dof_values(src);
    values_all(phi_m.get_data_ptr(), true, dim - 1);
    face_flux(phi_m, phi_m.get_data_ptr(), nullptr, nullptr, true);
    values_all(phi_m.get_data_ptr(), false, dim - 1);
    phi_m.distribute_local_to_global(dst);
  }
  void local_apply_inverse_mass_matrix(const MF &, Vec &dst, const Vec &src,
                                       const ID cell)
  {
    FEEvaluationInverse<dim_x, dim_v> phi(mf);
    phi.reinit(cell);
    phi.read_dof_values(src);
    inverse_mass(phi, phi.get_data_ptr());
    phi.set_dof_values(dst);
  }

  // element-centric: the cell's own faces inside the cell operation
  void local_apply_cell_centric(const MF &, Vec &dst, const Vec &src,
                                const ID cell)
  {
    FEEvaluationInverse<dim_x, dim_v> phi(mf);
    std::vector<double> buffer(nq), tempp((size_t)nq * dim);
    phi.reinit(cell);
    phi.read_dof_values(src);
    double *data = phi.get_data_ptr();
    values_all(data, true, dim);
    cell_terms(phi, data, buffer, tempp);
    values_all(data, false, dim);
    FEFaceEvaluation<dim_x, dim_v> phi_m(mf, true), phi_p(mf, false);
    for (unsigned int face = 0; face < 2 * dim; ++face)
      {
        phi_m.reinit(cell, face);
        phi_m.read_dof_values(src);
        values_all(phi_m.get_data_ptr(), true, dim - 1);
        const bool boundary = phi_m.at_boundary();
        if (!boundary)
          {
            phi_p.reinit(cell, face);
            phi_p.read_dof_values(src);
            values_all(phi_p.get_data_ptr(), true, dim - 1);
          }
        face_flux(phi_m, phi_m.get_data_ptr(), nullptr,
                  boundary ? nullptr : phi_p.get_data_ptr(), boundary);
        values_all(phi_m.get_data_ptr(), false, dim - 1);
        phi_m.distribute_to_buffer(data);
      }
    inverse_mass(phi, data);
    phi.set_dof_values(dst);
  }
};

} // namespace advection_host
