// The advection operator written against the host loop / evaluator
// interface (include/hyperdeal_amd_host.hpp) in hyper.deal's own shape
// (advection_op.hpp), run face-centric (MatrixFree::AdditionalData::use_ecl
// = false: loop with cell / face / boundary member functions, then the
// inverse mass matrix through cell_loop) and element-centric (use_ecl =
// true: loop_cell_centric), each compared with the oracle's apply of the
// same problem (test infrastructure: this program links the oracle library).
// The element-centric run also goes through a std::function and checks the
// loops' Timers entries.
//
// usage: advection_host <dim_x> <dim_v> <degree> <n_points> <collocation 0|1>
//                       <periodic 0|1> <skew>
// prints: "fcl <rel err> ecl <rel err> fn <rel err>"
#include "advection_op.hpp"
#include "hd_oracle.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace hyperdeal_amd::host;

template <int dim_x, int dim_v, int degree, int n_points>
int run(bool collocation, bool periodic, double skew)
{
  constexpr int dim = dim_x + dim_v;
  using Number = double;
  using VA = VectorizedArray<Number>;
  using VectorType = std::vector<Number>;
  using MF = MatrixFree<dim_x, dim_v, Number, VA>;
  using Op = advection_host::AdvectionOperation<dim_x, dim_v, degree,
                                                n_points, Number, VectorType,
                                                VA>;
  const int cells[6] = {3, 2, 3, 2, 2, 3};
  const double vel[6] = {0.7, -0.45, 0.3, -0.8, 0.55, -0.25};
  std::array<int, dim_x> ncx{};
  std::array<int, dim_v> ncv{};
  std::array<double, dim_x> lox{}, hix{};
  std::array<double, dim_v> lov{}, hiv{};
  std::array<double, dim> a{};
  int n_cells[6];
  double lo6[6], hi6[6], a6[6];
  for (int d = 0; d < dim; ++d)
    {
      n_cells[d] = cells[d];
      lo6[d] = -0.25 * d;
      hi6[d] = lo6[d] + 0.6 + 0.2 * d;
      a6[d] = a[d] = vel[d];
      if (d < dim_x)
        {
          ncx[d] = cells[d];
          lox[d] = lo6[d];
          hix[d] = hi6[d];
        }
      else
        {
          ncv[d - dim_x] = cells[d];
          lov[d - dim_x] = lo6[d];
          hiv[d - dim_x] = hi6[d];
        }
    }
  // the two low-dimensional spaces with the reference's quadrature list
  CartesianMatrixFree<dim_x> mfx;
  CartesianMatrixFree<dim_v> mfv;
  const auto quads = advection_quadratures(degree, n_points, collocation);
  mfx.reinit(ncx, lox, hix, periodic, degree, quads);
  mfv.reinit(ncv, lov, hiv, periodic, degree, quads);

  VectorType src, dst_fcl, dst_ecl, dst_fn, ref;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1.0, 1.0);

  // face-centric
  MF data_fcl(mfx, mfv);
  {
    typename MF::AdditionalData ad;
    ad.use_ecl = false;
    data_fcl.reinit(ad);
  }
  data_fcl.initialize_dof_vector(src);
  data_fcl.initialize_dof_vector(dst_fcl);
  for (auto &x : src)
    x = U(rng);
  Timers timers;
  Op op_fcl(data_fcl);
  op_fcl.reinit(a, skew);
  op_fcl.apply(dst_fcl, src, 0.0, &timers);

  // element-centric
  MF data_ecl(mfx, mfv);
  data_ecl.reinit();
  data_ecl.initialize_dof_vector(dst_ecl);
  Op op_ecl(data_ecl);
  op_ecl.reinit(a, skew);
  op_ecl.apply(dst_ecl, src, 0.0, &timers);

  // the std::function form of loop_cell_centric
  data_ecl.initialize_dof_vector(dst_fn);
  data_ecl.template loop_cell_centric<VectorType, VectorType>(
    [&op_ecl](const MF &m, VectorType &d, const VectorType &s,
              const TensorID id) {
      op_ecl.local_apply_advect_and_inverse_mass_matrix(m, d, s, id);
    },
    dst_fn, src, MF::DataAccessOnFaces::values, nullptr);

  // the loops timed what the reference's loops time
  const char *labels[] = {"FCL:advection:cell_loop", "FCL:advection:face_loop_x",
                          "FCL:mass", "ECL:loop"};
  for (const char *l : labels)
    if (!timers.find(l) || timers.find(l)->get_counter() == 0)
      {
        std::printf("missing timer %s\n", l);
        return 1;
      }

  // the function parameters: one per direction (HDO_FUNC_ZERO ignores them)
  const double fp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  hdo_problem *p =
    hdo_create(dim_x, dim_v, degree, n_points, collocation ? 1 : 0, n_cells,
               lo6, hi6, periodic ? 1 : 0, periodic ? 1 : 0, skew, a6,
               HDO_BC_HOMOGENEOUS, HDO_FUNC_ZERO, fp);
  if (!p || hdo_n_dofs(p) != (long long)data_ecl.n_dofs())
    {
      std::printf("oracle setup failed\n");
      return 1;
    }
  ref.assign(src.size(), 0.0);
  hdo_apply(p, src.data(), ref.data(), 0.0, 1);
  double e1 = 0.0, e2 = 0.0, e3 = 0.0, m = 0.0;
  for (size_t i = 0; i < ref.size(); ++i)
    {
      e1 = std::max(e1, std::fabs(dst_fcl[i] - ref[i]));
      e2 = std::max(e2, std::fabs(dst_ecl[i] - ref[i]));
      e3 = std::max(e3, std::fabs(dst_fn[i] - ref[i]));
      m = std::max(m, std::fabs(ref[i]));
    }
  hdo_destroy(p);
  std::printf("fcl %.3e ecl %.3e fn %.3e\n", e1 / m, e2 / m, e3 / m);
  return 0;
}

int main(int argc, char **argv)
{
  if (argc < 8)
    {
      std::printf("usage: %s dim_x dim_v degree n_points collocation "
                  "periodic skew\n",
                  argv[0]);
      return 2;
    }
  const int dx = std::atoi(argv[1]), dv = std::atoi(argv[2]);
  const int k = std::atoi(argv[3]), nq = std::atoi(argv[4]);
  const bool coll = std::atoi(argv[5]) != 0;
  const bool periodic = std::atoi(argv[6]) != 0;
  const double skew = std::atof(argv[7]);
#define HD_CASE(DX, DV, K, NQ)                                               \
  if (dx == DX && dv == DV && k == K && nq == NQ)                           \
    return run<DX, DV, K, NQ>(coll, periodic, skew);
  HD_CASE(1, 1, 1, 2) HD_CASE(1, 1, 2, 3) HD_CASE(1, 1, 3, 4)
  HD_CASE(1, 1, 3, 5) HD_CASE(1, 1, 2, 4)
  HD_CASE(1, 2, 3, 4) HD_CASE(1, 2, 2, 4)
  HD_CASE(2, 1, 1, 2) HD_CASE(2, 1, 1, 3)
  HD_CASE(2, 2, 1, 2) HD_CASE(2, 2, 2, 3) HD_CASE(2, 2, 3, 4)
  HD_CASE(2, 2, 3, 5) HD_CASE(2, 2, 2, 4)
  HD_CASE(3, 3, 1, 2) HD_CASE(3, 3, 2, 3) HD_CASE(3, 3, 1, 3)
#undef HD_CASE
  std::printf("unsupported case\n");
  return 2;
}
