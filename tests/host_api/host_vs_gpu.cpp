// The advection operator written against the host loop / evaluator
// interface (advection_op.hpp: hyper.deal's AdvectionOperation shape,
// face-centric and element-centric) against the GPU operator of the same
// problem through the C++ wrapper of the C ABI (include/hyperdeal_amd.hpp):
// the two sides of the drop-in boundary give the same apply.
//
// usage: host_vs_gpu <dim_x> <dim_v> <degree> <n_points> <collocation 0|1>
//                    <periodic 0|1> <skew>
// prints: "fcl <rel err> ecl <rel err>" (each: GPU against the host loop)
#include "advection_op.hpp"
#include "hyperdeal_amd.hpp"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

namespace hh = hyperdeal_amd::host;

template <int dim_x, int dim_v, int degree, int n_points>
int run(bool collocation, bool periodic, double skew)
{
  constexpr int dim = dim_x + dim_v;
  using VectorType = std::vector<double>;
  using MF = hh::MatrixFree<dim_x, dim_v, double, hh::VectorizedArray<double>>;
  using Op = advection_host::AdvectionOperation<dim_x, dim_v, degree,
                                                n_points, double, VectorType,
                                                double>;
  std::array<int, dim> nc{};
  std::array<double, dim> lo{}, hi{}, a{};
  std::array<int, dim_x> ncx{};
  std::array<int, dim_v> ncv{};
  std::array<double, dim_x> lox{}, hix{};
  std::array<double, dim_v> lov{}, hiv{};
  const int cells[6] = {4, 3, 3, 2, 3, 2};
  const double vel[6] = {-0.6, 0.35, 0.8, -0.45, 0.25, -0.9};
  for (int d = 0; d < dim; ++d)
    {
      nc[d] = cells[d];
      lo[d] = 0.1 * d;
      hi[d] = lo[d] + 0.5 + 0.25 * d;
      a[d] = vel[d];
      if (d < dim_x)
        {
          ncx[d] = nc[d];
          lox[d] = lo[d];
          hix[d] = hi[d];
        }
      else
        {
          ncv[d - dim_x] = nc[d];
          lov[d - dim_x] = lo[d];
          hiv[d - dim_x] = hi[d];
        }
    }
  hh::CartesianMatrixFree<dim_x> mfx;
  hh::CartesianMatrixFree<dim_v> mfv;
  const auto quads = hh::advection_quadratures(degree, n_points, collocation);
  mfx.reinit(ncx, lox, hix, periodic, degree, quads);
  mfv.reinit(ncv, lov, hiv, periodic, degree, quads);
  VectorType src, host_fcl, host_ecl;
  std::array<VectorType *, 2> out = {&host_fcl, &host_ecl};
  for (int ecl = 0; ecl < 2; ++ecl)
    {
      MF data(mfx, mfv);
      typename MF::AdditionalData ad;
      ad.use_ecl = ecl == 1;
      data.reinit(ad);
      if (src.empty())
        {
          data.initialize_dof_vector(src);
          std::mt19937_64 rng(11);
          std::uniform_real_distribution<double> U(-1.0, 1.0);
          for (auto &x : src)
            x = U(rng);
        }
      data.initialize_dof_vector(*out[ecl]);
      Op op(data);
      op.reinit(a, skew);
      op.apply(*out[ecl], src, 0.0);
    }

  typename hyperdeal_amd::MatrixFree<dim_x, dim_v>::AdditionalData gad;
  gad.n_cells = nc;
  gad.lo = lo;
  gad.hi = hi;
  gad.periodic_x = periodic;
  gad.periodic_v = periodic;
  gad.degree = degree;
  gad.n_points = n_points;
  gad.do_collocation = collocation;
  hyperdeal_amd::MatrixFree<dim_x, dim_v> gmf;
  gmf.reinit(gad);
  hyperdeal_amd::AdvectionOperation<dim_x, dim_v, double> gop(gmf);
  gop.reinit(a, skew, HD_BC_HOMOGENEOUS, HD_FUNC_ZERO);
  hyperdeal_amd::DeviceVector<double> s, d;
  gmf.initialize_dof_vector(s);
  gmf.initialize_dof_vector(d);
  s.copy_from_host(src);
  gop.apply(d, s, 0.0);
  const std::vector<double> g = d.to_host();
  double e[2] = {0.0, 0.0}, m = 0.0;
  for (size_t i = 0; i < g.size(); ++i)
    {
      e[0] = std::max(e[0], std::fabs(g[i] - host_fcl[i]));
      e[1] = std::max(e[1], std::fabs(g[i] - host_ecl[i]));
      m = std::max(m, std::fabs(host_ecl[i]));
    }
  std::printf("fcl %.3e ecl %.3e\n", e[0] / m, e[1] / m);
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
  try
    {
#define HD_CASE(DX, DV, K, NQ)                                               \
  if (dx == DX && dv == DV && k == K && nq == NQ)                           \
    return run<DX, DV, K, NQ>(coll, periodic, skew);
      HD_CASE(1, 1, 3, 4) HD_CASE(1, 1, 3, 5)
      HD_CASE(2, 2, 2, 3) HD_CASE(2, 2, 3, 4) HD_CASE(2, 2, 3, 5)
      HD_CASE(3, 3, 1, 2) HD_CASE(3, 3, 2, 3) HD_CASE(3, 3, 1, 3)
#undef HD_CASE
    }
  catch (const std::exception &ex)
    {
      std::printf("error: %s\n", ex.what());
      return 1;
    }
  std::printf("unsupported case\n");
  return 2;
}
