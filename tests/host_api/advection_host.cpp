// The advection operator written against the host loop / evaluator
// interface (include/hyperdeal_amd_host.hpp), the way hyper.deal writes it:
// face-centric (cell, face and boundary operations through
// MatrixFree::loop, then the inverse mass matrix through cell_loop;
// include/hyper.deal/operators/advection/advection_operation.h:573-1131)
// and element-centric (every face inside the cell operation through
// loop_cell_centric; advection_operation.h:221-566).  Both are compared
// with the oracle's apply of the same problem (test infrastructure: this
// program links the oracle library).
//
// usage: advection_host <dim_x> <dim_v> <degree> <periodic 0|1> <skew>
// prints: "fcl <rel err> ecl <rel err>"
#include "hyperdeal_amd_host.hpp"

#include "hd_oracle.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

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

template <int dim_x, int dim_v>
int run(int degree, bool periodic, double skew)
{
  constexpr int dim = dim_x + dim_v;
  std::array<int, dim> nc{};
  std::array<double, dim> lo{}, hi{}, a{};
  const int cells[6] = {3, 2, 3, 2, 2, 3};
  const double vel[6] = {0.7, -0.45, 0.3, -0.8, 0.55, -0.25};
  for (int d = 0; d < dim; ++d)
    {
      nc[d] = cells[d];
      lo[d] = -0.25 * d;
      hi[d] = lo[d] + 0.6 + 0.2 * d;
      a[d] = vel[d];
    }
  MatrixFree<dim_x, dim_v> mf;
  mf.reinit(nc, lo, hi, degree, periodic, periodic);
  Advection<dim_x, dim_v> op(mf, a, skew);

  std::vector<double> src, dst, dst2, ref(mf.n_dofs());
  mf.initialize_dof_vector(src);
  mf.initialize_dof_vector(dst);
  mf.initialize_dof_vector(dst2);
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  for (auto &x : src)
    x = U(rng);

  // face-centric, member-pointer form, then the inverse mass matrix
  using A = Advection<dim_x, dim_v>;
  mf.loop(&A::local_apply_cell, &A::local_apply_face,
          &A::local_apply_boundary, &op, dst, src);
  mf.cell_loop(&A::local_apply_inverse_mass_matrix, &op, dst, dst);
  // element-centric, std::function form
  mf.template loop_cell_centric<std::vector<double>, std::vector<double>>(
    [&op](const MatrixFree<dim_x, dim_v> &m, std::vector<double> &d,
          const std::vector<double> &s, const TensorID id) {
      op.local_apply_cell_centric(m, d, s, id);
    },
    dst2, src);

  int n_cells[6];
  double lo6[6], hi6[6], a6[6];
  for (int d = 0; d < dim; ++d)
    {
      n_cells[d] = nc[d];
      lo6[d] = lo[d];
      hi6[d] = hi[d];
      a6[d] = a[d];
    }
  const double fp[1] = {0.0};
  hdo_problem *p =
    hdo_create(dim_x, dim_v, degree, degree + 1, 0, n_cells, lo6, hi6,
               periodic ? 1 : 0, periodic ? 1 : 0, skew, a6,
               HDO_BC_HOMOGENEOUS, HDO_FUNC_ZERO, fp);
  if (!p || hdo_n_dofs(p) != (long long)mf.n_dofs())
    {
      std::printf("oracle setup failed\n");
      return 1;
    }
  hdo_apply(p, src.data(), ref.data(), 0.0, 1);
  double e1 = 0.0, e2 = 0.0, m = 0.0;
  for (size_t i = 0; i < ref.size(); ++i)
    {
      e1 = std::max(e1, std::fabs(dst[i] - ref[i]));
      e2 = std::max(e2, std::fabs(dst2[i] - ref[i]));
      m = std::max(m, std::fabs(ref[i]));
    }
  hdo_destroy(p);
  std::printf("fcl %.3e ecl %.3e\n", e1 / m, e2 / m);
  return 0;
}

int main(int argc, char **argv)
{
  if (argc < 6)
    {
      std::printf("usage: %s dim_x dim_v degree periodic skew\n", argv[0]);
      return 2;
    }
  const int dx = std::atoi(argv[1]), dv = std::atoi(argv[2]);
  const int k = std::atoi(argv[3]);
  const bool periodic = std::atoi(argv[4]) != 0;
  const double skew = std::atof(argv[5]);
  if (dx == 1 && dv == 1)
    return run<1, 1>(k, periodic, skew);
  if (dx == 1 && dv == 2)
    return run<1, 2>(k, periodic, skew);
  if (dx == 2 && dv == 2)
    return run<2, 2>(k, periodic, skew);
  if (dx == 2 && dv == 1)
    return run<2, 1>(k, periodic, skew);
  if (dx == 3 && dv == 3)
    return run<3, 3>(k, periodic, skew);
  std::printf("unsupported dims\n");
  return 2;
}
