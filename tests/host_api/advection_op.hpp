// The advection operator written against the host loop / evaluator
// interface (include/hyperdeal_amd_host.hpp) in the shape of hyper.deal's
// AdvectionOperation (include/hyper.deal/operators/advection/
// advection_operation.h): the same class template parameters, evaluators
// built in reinit() with (data, 0, 0, 0, 0) and the inverse-mass evaluator on
// quadrature 1 (Gauss, degree + 1 points), apply() calling
//   data.loop(&This::local_apply_cell, &This::local_apply_face,
//             &This::local_apply_boundary, this, dst, src,
//             DataAccessOnFaces::values, DataAccessOnFaces::values, timers)
//   + data.cell_loop(&This::local_apply_inverse_mass_matrix, this, dst, dst)
// for face-centric loops and
//   data.loop_cell_centric(&This::local_apply_advect_and_inverse_mass_matrix,
//                          this, dst, src, DataAccessOnFaces::values, timers)
// for element-centric ones (advection_operation.h:134-209), with the
// over-integration (n_points != degree + 1: basis change) and collocation
// (Gauss-Lobatto points: no S) branches.  Constant velocity a, skew factor,
// homogeneous Dirichlet data (u+ = -u-) on boundary faces.
#pragma once

#include "hyperdeal_amd_host.hpp"

#include <array>
#include <cmath>
#include <memory>
#include <vector>

namespace advection_host
{
using namespace hyperdeal_amd::host;

template <int dim_x, int dim_v, int degree, int n_points, typename Number,
          typename VectorType, typename VectorizedArrayType>
class AdvectionOperation
{
public:
  using This = AdvectionOperation<dim_x, dim_v, degree, n_points, Number,
                                  VectorType, VectorizedArrayType>;
  using VNumber = VectorizedArrayType;
  using MF = MatrixFree<dim_x, dim_v, Number, VectorizedArrayType>;
  using ID = typename MF::ID;
  static constexpr int dim = dim_x + dim_v;
  static constexpr unsigned int nq_all = detail::cpow(n_points, dim);

  using FECellEval =
    FEEvaluation<dim_x, dim_v, degree, n_points, Number, VNumber>;
  using FEFaceEval =
    FEFaceEvaluation<dim_x, dim_v, degree, n_points, Number, VNumber>;
  using FECellEval_inv =
    FEEvaluationInverse<dim_x, dim_v, degree, n_points, Number, VNumber>;
  using FECellEval_inv_co =
    FEEvaluationInverse<dim_x, dim_v, degree, degree + 1, Number, VNumber>;

  explicit AdvectionOperation(const MF &data) : data(data) {}

  // advection_operation.h:97-129
  void reinit(const std::array<double, dim> &velocity, double factor_skew)
  {
    a = velocity;
    this->factor_skew = factor_skew;
    do_collocation =
      data.get_matrix_free_x().get_shape_info(0, 0).data[0].element_type ==
      ElementType::tensor_symmetric_collocation;
    phi_cell.reset(new FECellEval(data, 0, 0, 0, 0));
    phi_cell_inv.reset(new FECellEval_inv(data, 0, 0, 0, 0));
    phi_cell_inv_co.reset(new FECellEval_inv_co(
      data, 0, 0, do_collocation ? 0 : 1, do_collocation ? 0 : 1));
    phi_face_m.reset(new FEFaceEval(data, true, 0, 0, 0, 0));
    phi_face_p.reset(new FEFaceEval(data, false, 0, 0, 0, 0));
  }

  // advection_operation.h:134-209
  void apply(VectorType &dst, const VectorType &src, const Number time,
             Timers *timers = nullptr)
  {
    (void)time;
    if (!data.is_ecl_supported()) // FCL
      {
        if (timers != nullptr)
          timers->enter("FCL");
        {
          ScopedTimerWrapper timer(timers, "advection");
          if (timers != nullptr)
            timers->enter("advection");
          data.loop(&This::local_apply_cell, &This::local_apply_face,
                    &This::local_apply_boundary, this, dst, src,
                    MF::DataAccessOnFaces::values,
                    MF::DataAccessOnFaces::values, timers);
          if (timers != nullptr)
            timers->leave();
        }
        {
          ScopedTimerWrapper timer(timers, "mass");
          data.cell_loop(&This::local_apply_inverse_mass_matrix, this, dst,
                         dst);
        }
        if (timers != nullptr)
          timers->leave();
      }
    else // ECL
      {
        if (timers != nullptr)
          timers->enter("ECL");
        data.loop_cell_centric(
          &This::local_apply_advect_and_inverse_mass_matrix, this, dst, src,
          MF::DataAccessOnFaces::values, timers);
        if (timers != nullptr)
          timers->leave();
      }
  }

private:
  // nodal values -> point values (or back, transposed) in dims directions
  template <int dims>
  void to_points(VNumber *p) const
  {
    if (do_collocation)
      return;
    const auto &S = *phi_cell->get_shape_values();
    if constexpr (degree + 1 == n_points)
      {
        const EvaluatorTensorProduct<dims, degree + 1, n_points, VNumber> ev(
          S, {}, {});
        static_for_dirs<0, dims, true>(ev, p);
      }
    else
      FEEvaluationImplBasisChange<dims, degree + 1, n_points,
                                  VNumber>::do_forward(1, S, p, p);
  }
  template <int dims>
  void from_points(VNumber *p) const
  {
    if (do_collocation)
      return;
    const auto &S = *phi_cell->get_shape_values();
    if constexpr (degree + 1 == n_points)
      {
        const EvaluatorTensorProduct<dims, degree + 1, n_points, VNumber> ev(
          S, {}, {});
        static_for_dirs<dims - 1, dims, false>(ev, p);
      }
    else
      FEEvaluationImplBasisChange<dims, degree + 1, n_points,
                                  VNumber>::do_backward(1, S, false, p, p);
  }
  // values<d, forward, false> for d = first, first +- 1, ... within [0, dims)
  template <int d, int dims, bool forward, typename E>
  static void static_for_dirs(const E &ev, VNumber *p)
  {
    if constexpr (d >= 0 && d < dims)
      {
        ev.template values<d, forward, false>(p, p);
        static_for_dirs<forward ? d + 1 : d - 1, dims, forward>(ev, p);
      }
  }

  // the cell integrals at the points (advection_operation.h:316-403 /
  // 680-810): data holds the point values on entry and the integrated
  // test-function weights on exit; buffer: a copy of the point values
  void cell_integrals(FECellEval &phi, VNumber *data_ptr,
                      const VNumber *buffer) const
  {
    const auto &D = *phi.get_shape_gradients();
    const EvaluatorTensorProduct<dim, n_points, n_points, VNumber> eval_(
      {}, D, {});
    std::vector<VNumber> scratch((size_t)nq_all * dim);
    VNumber *tempp = scratch.data();
    constexpr unsigned int nqx = detail::cpow(n_points, dim_x);
    constexpr unsigned int nqv = detail::cpow(n_points, dim_v);
    // x-space
    if (factor_skew != 0.0)
      grad_dirs<0, dim_x, true>(eval_, buffer, tempp);
    for (unsigned int qv = 0, q = 0; qv < nqv; ++qv)
      for (unsigned int qx = 0; qx < nqx; ++qx, ++q)
        {
          if (factor_skew != 0.0)
            {
              const auto g = phi.get_gradient_x(tempp, q, qx, qv);
              VNumber gv = 0.0;
              for (int d = 0; d < dim_x; ++d)
                gv += g[d] * a[d];
              phi.template submit_value<false>(data_ptr, -factor_skew * gv, q,
                                               qx, qv);
            }
          if (factor_skew != 1.0)
            {
              std::array<VNumber, dim_x> gi{};
              for (int d = 0; d < dim_x; ++d)
                gi[d] = (1.0 - factor_skew) * buffer[q] * a[d];
              phi.submit_gradient_x(tempp, gi.data(), q, qx, qv);
            }
        }
    if (factor_skew != 1.0)
      grad_dirs_t<0, dim_x>(eval_, tempp, data_ptr, factor_skew != 0.0);
    // v-space
    if (factor_skew != 0.0)
      grad_dirs<dim_x, dim, true>(eval_, buffer, tempp);
    for (unsigned int qv = 0, q = 0; qv < nqv; ++qv)
      for (unsigned int qx = 0; qx < nqx; ++qx, ++q)
        {
          if (factor_skew != 0.0)
            {
              const auto g = phi.get_gradient_v(tempp, q, qx, qv);
              VNumber gv = 0.0;
              for (int d = 0; d < dim_v; ++d)
                gv += g[d] * a[dim_x + d];
              phi.template submit_value<true>(data_ptr, -factor_skew * gv, q,
                                              qx, qv);
            }
          if (factor_skew != 1.0)
            {
              std::array<VNumber, dim_v> gi{};
              for (int d = 0; d < dim_v; ++d)
                gi[d] = (1.0 - factor_skew) * buffer[q] * a[dim_x + d];
              phi.submit_gradient_v(tempp, gi.data(), q, qx, qv);
            }
        }
    if (factor_skew != 1.0)
      grad_dirs_t<dim_x, dim>(eval_, tempp, data_ptr, true);
  }
  // tempp[(d - d0) nq ...] = D along d of in, d in [d0, d1)
  template <int d0, int d1, bool fw, typename E>
  static void grad_dirs(const E &ev, const VNumber *in, VNumber *tempp)
  {
    if constexpr (d0 < d1)
      {
        constexpr int e = d0 < dim_x ? d0 : d0 - dim_x;
        ev.template gradients<d0, fw, false>(in, tempp + (size_t)e * nq_all);
        grad_dirs<d0 + 1, d1, fw>(ev, in, tempp);
      }
  }
  // out (+)= sum_d D^T along d of tempp[(d - d0) nq ...]; the first
  // direction adds only when add_first
  template <int d0, int d1, typename E>
  static void grad_dirs_t(const E &ev, const VNumber *tempp, VNumber *out,
                          bool add_first)
  {
    if constexpr (d0 < d1)
      {
        constexpr int e = d0 < dim_x ? d0 : d0 - dim_x;
        const bool first = d0 == 0 || d0 == dim_x;
        if (first && !add_first)
          ev.template gradients<d0, false, false>(tempp + (size_t)e * nq_all,
                                                  out);
        else
          ev.template gradients<d0, false, true>(tempp + (size_t)e * nq_all,
                                                 out);
        grad_dirs_t<d0 + 1, d1>(ev, tempp, out, true);
      }
  }

  // upwind flux at the face points (advection_operation.h:455-520):
  // submits to the minus side (and the plus side when given)
  void flux(FEFaceEval &phi_m, VNumber *d1, FEFaceEval *phi_p, VNumber *d2,
            bool boundary) const
  {
    const bool in_x = phi_m.get_type() == ID::SpaceType::X;
    const unsigned int nqx = phi_m.face_points_x();
    const unsigned int nqv = detail::cpow(n_points, dim - 1) / nqx;
    for (unsigned int qv = 0, q = 0; qv < nqv; ++qv)
      for (unsigned int qx = 0; qx < nqx; ++qx, ++q)
        {
          const VNumber u_minus = d1[q];
          const VNumber u_plus = boundary ? -u_minus : d2[q];
          VNumber ns = 0.0;
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
          const VNumber f = 0.5 * ((u_minus + u_plus) * ns +
                                   std::abs(ns) * (u_minus - u_plus)) *
                            alpha;
          if (in_x)
            phi_m.template submit_value<ID::SpaceType::X>(
              d1, f - factor_skew * u_minus * ns, q, qx, qv);
          else
            phi_m.template submit_value<ID::SpaceType::V>(
              d1, f - factor_skew * u_minus * ns, q, qx, qv);
          if (phi_p)
            {
              if (in_x)
                phi_p->template submit_value<ID::SpaceType::X>(
                  d2, -f + factor_skew * u_plus * ns, q, qx, qv);
              else
                phi_p->template submit_value<ID::SpaceType::V>(
                  d2, -f + factor_skew * u_plus * ns, q, qx, qv);
            }
        }
  }

  // 1/JxW and the inverse shape values along every direction
  // (advection_operation.h:528-559)
  void inverse_mass_at_points(FECellEval_inv &phi_inv, VNumber *data_ptr) const
  {
    constexpr unsigned int nqx = detail::cpow(n_points, dim_x);
    constexpr unsigned int nqv = detail::cpow(n_points, dim_v);
    for (unsigned int qv = 0, q = 0; qv < nqv; ++qv)
      for (unsigned int qx = 0; qx < nqx; ++qx, ++q)
        phi_inv.submit_inv(data_ptr, q, qx, qv);
    if (do_collocation)
      return;
    const auto &P = *phi_inv.get_inverse_shape();
    if constexpr (degree + 1 == n_points)
      {
        const EvaluatorTensorProduct<dim, degree + 1, n_points, VNumber> ev(
          {}, {}, P);
        hess_dirs<dim - 1>(ev, data_ptr);
      }
    else
      FEEvaluationImplBasisChange<dim, degree + 1, n_points,
                                  VNumber>::do_backward(1, P, false, data_ptr,
                                                        data_ptr);
  }
  template <int d, typename E>
  static void hess_dirs(const E &ev, VNumber *p)
  {
    if constexpr (d >= 0)
      {
        ev.template hessians<d, false, false>(p, p);
        hess_dirs<d - 1>(ev, p);
      }
  }

public:
  // ECL: advection and inverse mass matrix of one cell
  // (advection_operation.h:221-566)
  void local_apply_advect_and_inverse_mass_matrix(const MF &data,
                                                  VectorType &dst,
                                                  const VectorType &src,
                                                  const ID cell)
  {
    auto &phi = *phi_cell;
    auto &phi_m = *phi_face_m;
    auto &phi_p = *phi_face_p;
    auto &phi_inv = *phi_cell_inv;
    VNumber *data_ptr = phi.get_data_ptr();
    VNumber *data_ptr1 = phi_m.get_data_ptr();
    VNumber *data_ptr2 = phi_p.get_data_ptr();
    VNumber *buffer = phi_inv.get_data_ptr();

    // 1) cell contribution
    phi.reinit(cell);
    phi.read_dof_values(src);
    to_points<dim>(data_ptr);
    for (unsigned int i = 0; i < nq_all; ++i)
      buffer[i] = data_ptr[i];
    cell_integrals(phi, data_ptr, buffer);

    // 2) faces
    for (unsigned int face = 0; face < 2 * dim; ++face)
      {
        phi_m.reinit(cell, face);
        const auto bid = data.get_faces_by_cells_boundary_id(cell, face);
        const bool inner = bid == internal_face_boundary_id;
        if (inner)
          {
            phi_p.reinit(cell, face);
            phi_p.read_dof_values(src);
          }
        if (!do_collocation)
          {
            FEFaceNormalEvaluationImpl<dim_x, dim_v, n_points - 1, Number>::
              template interpolate_quadrature<true, false>(
                1, EvaluationFlags::values,
                data.get_matrix_free_x().get_shape_info(), buffer, data_ptr1,
                face);
            if (inner)
              to_points<dim - 1>(data_ptr2);
          }
        else
          phi_m.read_dof_values_from_buffer(buffer);
        flux(phi_m, data_ptr1, nullptr, data_ptr2, !inner);
        if (!do_collocation)
          FEFaceNormalEvaluationImpl<dim_x, dim_v, n_points - 1, Number>::
            template interpolate_quadrature<false, true>(
              1, EvaluationFlags::values,
              data.get_matrix_free_x().get_shape_info(), data_ptr1, data_ptr,
              face);
        else
          phi_m.distribute_to_buffer(data_ptr);
      }

    // 3) inverse mass matrix
    phi_inv.reinit(cell);
    inverse_mass_at_points(phi_inv, data_ptr);
    phi.set_dof_values(dst);
  }

  // FCL: cell integrals (advection_operation.h:639-818)
  void local_apply_cell(const MF &, VectorType &dst, const VectorType &src,
                        const ID cell)
  {
    auto &phi = *phi_cell;
    VNumber *data_ptr = phi.get_data_ptr();
    phi.reinit(cell);
    phi.read_dof_values(src);
    to_points<dim>(data_ptr);
    std::vector<VNumber> buffer(data_ptr, data_ptr + nq_all);
    cell_integrals(phi, data_ptr, buffer.data());
    from_points<dim>(data_ptr);
    phi.set_dof_values(dst);
  }

  // FCL: inner faces, both sides (advection_operation.h:825-1000)
  void local_apply_face(const MF &, VectorType &dst, const VectorType &src,
                        const ID face)
  {
    auto &phi_m = *phi_face_m;
    auto &phi_p = *phi_face_p;
    VNumber *data_ptr1 = phi_m.get_data_ptr();
    VNumber *data_ptr2 = phi_p.get_data_ptr();
    phi_m.reinit(face);
    phi_p.reinit(face);
    phi_m.read_dof_values(src);
    to_points<dim - 1>(data_ptr1);
    phi_p.read_dof_values(src);
    to_points<dim - 1>(data_ptr2);
    flux(phi_m, data_ptr1, &phi_p, data_ptr2, false);
    from_points<dim - 1>(data_ptr1);
    phi_m.distribute_local_to_global(dst);
    from_points<dim - 1>(data_ptr2);
    phi_p.distribute_local_to_global(dst);
  }

  // FCL: boundary faces (advection_operation.h:1009-1131)
  void local_apply_boundary(const MF &, VectorType &dst, const VectorType &src,
                            const ID face)
  {
    auto &phi_m = *phi_face_m;
    VNumber *data_ptr1 = phi_m.get_data_ptr();
    phi_m.reinit(face);
    phi_m.read_dof_values(src);
    to_points<dim - 1>(data_ptr1);
    flux(phi_m, data_ptr1, nullptr, nullptr, true);
    from_points<dim - 1>(data_ptr1);
    phi_m.distribute_local_to_global(dst);
  }

  // FCL: inverse mass matrix on Gauss degree + 1 points
  // (advection_operation.h:573-632)
  void local_apply_inverse_mass_matrix(const MF &, VectorType &dst,
                                       const VectorType &src, const ID cell)
  {
    auto &phi_inv = *phi_cell_inv_co;
    VNumber *data_ptr = phi_inv.get_data_ptr();
    phi_inv.reinit(cell);
    phi_inv.read_dof_values(src);
    const EvaluatorTensorProduct<dim, degree + 1, degree + 1, VNumber>
      eval_inv({}, {}, *phi_inv.get_inverse_shape());
    if (!do_collocation)
      inv_dirs<0, true>(eval_inv, data_ptr);
    constexpr unsigned int nx = detail::cpow(degree + 1, dim_x);
    constexpr unsigned int nv = detail::cpow(degree + 1, dim_v);
    for (unsigned int qv = 0, q = 0; qv < nv; ++qv)
      for (unsigned int qx = 0; qx < nx; ++qx, ++q)
        phi_inv.submit_inv(data_ptr, q, qx, qv);
    if (!do_collocation)
      inv_dirs<dim - 1, false>(eval_inv, data_ptr);
    phi_inv.set_dof_values(dst);
  }

private:
  template <int d, bool fw, typename E>
  static void inv_dirs(const E &ev, VNumber *p)
  {
    if constexpr (d >= 0 && d < dim)
      {
        ev.template hessians<d, fw, false>(p, p);
        inv_dirs<fw ? d + 1 : d - 1, fw>(ev, p);
      }
  }

  const MF &data;
  std::array<double, dim> a{};
  double factor_skew = 0.0;
  const double alpha = 1.0;
  bool do_collocation = false;
  std::shared_ptr<FECellEval> phi_cell;
  std::shared_ptr<FECellEval_inv> phi_cell_inv;
  std::shared_ptr<FECellEval_inv_co> phi_cell_inv_co;
  std::shared_ptr<FEFaceEval> phi_face_m;
  std::shared_ptr<FEFaceEval> phi_face_p;
};

} // namespace advection_host
