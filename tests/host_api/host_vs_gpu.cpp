// The advection operator written against the host loop / evaluator
// interface against the GPU operator of the same problem through the C ABI
// (include/hyperdeal_amd.hpp): the two sides of the drop-in boundary give
// the same apply.
//
// usage: host_vs_gpu <dim_x> <dim_v> <degree> <periodic 0|1> <skew>
// prints: "gpu <rel err>"
#include "advection_op.hpp"
#include "hyperdeal_amd.hpp"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using advection_host::Advection;

template <int dim_x, int dim_v>
int run(int degree, bool periodic, double skew)
{
  constexpr int dim = dim_x + dim_v;
  std::array<int, dim> nc{};
  std::array<double, dim> lo{}, hi{}, a{};
  const int cells[6] = {4, 3, 3, 2, 3, 2};
  const double vel[6] = {-0.6, 0.35, 0.8, -0.45, 0.25, -0.9};
  for (int d = 0; d < dim; ++d)
    {
      nc[d] = cells[d];
      lo[d] = 0.1 * d;
      hi[d] = lo[d] + 0.5 + 0.25 * d;
      a[d] = vel[d];
    }
  hyperdeal_amd::host::MatrixFree<dim_x, dim_v> hmf;
  hmf.reinit(nc, lo, hi, degree, periodic, periodic);
  Advection<dim_x, dim_v> op(hmf, a, skew);
  std::vector<double> src, dst;
  hmf.initialize_dof_vector(src);
  hmf.initialize_dof_vector(dst);
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  for (auto &x : src)
    x = U(rng);
  using A = Advection<dim_x, dim_v>;
  hmf.loop(&A::local_apply_cell, &A::local_apply_face,
           &A::local_apply_boundary, &op, dst, src);
  hmf.cell_loop(&A::local_apply_inverse_mass_matrix, &op, dst, dst);

  typename hyperdeal_amd::MatrixFree<dim_x, dim_v>::AdditionalData ad;
  ad.n_cells = nc;
  ad.lo = lo;
  ad.hi = hi;
  ad.periodic_x = periodic;
  ad.periodic_v = periodic;
  ad.degree = degree;
  ad.n_points = degree + 1;
  hyperdeal_amd::MatrixFree<dim_x, dim_v> gmf;
  gmf.reinit(ad);
  hyperdeal_amd::AdvectionOperation<dim_x, dim_v, double> gop(gmf);
  gop.reinit(a, skew, HD_BC_HOMOGENEOUS, HD_FUNC_ZERO);
  hyperdeal_amd::DeviceVector<double> s, d;
  gmf.initialize_dof_vector(s);
  gmf.initialize_dof_vector(d);
  s.copy_from_host(src);
  gop.apply(d, s, 0.0);
  const std::vector<double> g = d.to_host();
  double e = 0.0, m = 0.0;
  for (size_t i = 0; i < g.size(); ++i)
    {
      e = std::max(e, std::fabs(g[i] - dst[i]));
      m = std::max(m, std::fabs(dst[i]));
    }
  std::printf("gpu %.3e\n", e / m);
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
  try
    {
      if (dx == 1 && dv == 1)
        return run<1, 1>(k, periodic, skew);
      if (dx == 2 && dv == 2)
        return run<2, 2>(k, periodic, skew);
      if (dx == 3 && dv == 3)
        return run<3, 3>(k, periodic, skew);
    }
  catch (const std::exception &ex)
    {
      std::printf("error: %s\n", ex.what());
      return 1;
    }
  std::printf("unsupported dims\n");
  return 2;
}
