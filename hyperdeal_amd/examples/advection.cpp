// advection.cpp -- the advection example (examples/advection/advection.cc,
// include/application.h) on the MI355X operator, through the C++ host
// interface: Cartesian box, initial condition at the nodes, CFL time step,
// LSRK time loop with the stages fused into the GPU apply, and the
// diagnostics rows of application.h:484-492.
//
//   advection dim_x dim_v degree n_points collocation periodic skew rk
//             cells_per_dir time_step final_time cfl tick [fp32]
//
// The case is the hyperrectangle of examples/advection/cases/
// hyperrectangle.h: box [-1,1]^D, velocity (1, 0.15, -0.05, 0, ...),
// f = sin(2 pi (x0 - t)) prod cos(2 pi (x_d - a_d t)).
#include <hyperdeal_amd.hpp>

#include <cstdio>
#include <cstdlib>
#include <string>

namespace hda = hyperdeal_amd;

template <int dim_x, int dim_v, typename Number>
int run(int argc, char **argv)
{
  constexpr int dim = dim_x + dim_v;
  const int degree = std::atoi(argv[3]);
  const int n_points = std::atoi(argv[4]);
  const bool collocation = std::atoi(argv[5]) != 0;
  const bool periodic = std::atoi(argv[6]) != 0;
  const double skew = std::atof(argv[7]);
  const std::string rk = argv[8];
  const int cells = std::atoi(argv[9]);
  const double time_step = std::atof(argv[10]);
  const double final_time = std::atof(argv[11]);
  const double cfl = std::atof(argv[12]);
  const double tick = std::atof(argv[13]);
  const double start_time = 0.0;

  hda::MatrixFree<dim_x, dim_v> mf;
  typename hda::MatrixFree<dim_x, dim_v>::AdditionalData ad;
  for (int d = 0; d < dim; ++d)
    {
      ad.n_cells[d] = cells;
      ad.lo[d] = -1.0;
      ad.hi[d] = 1.0;
    }
  ad.periodic_x = ad.periodic_v = periodic;
  ad.degree = degree;
  ad.n_points = n_points;
  ad.do_collocation = collocation;
  mf.reinit(ad);

  std::array<double, dim> a{};
  const double dirs[6] = {1.0, 0.15, -0.05, 0.0, 0.0, 0.0};
  for (int d = 0; d < dim; ++d)
    a[d] = dirs[d];
  hda::AdvectionOperation<dim_x, dim_v, Number> op(mf);
  op.reinit(a, skew, HD_BC_INHOMOGENEOUS, HD_FUNC_HYPERRECTANGLE);

  hda::DeviceVector<double> solution;
  hda::DeviceVector<Number> ki, ti;
  mf.initialize_dof_vector(solution);
  mf.initialize_dof_vector(ki);
  mf.initialize_dof_vector(ti);
  hda::VectorTools::interpolate(mf, HD_FUNC_HYPERRECTANGLE, a.data(),
                                start_time, solution);

  hda::TimeLoop loop;
  loop.start_time = start_time;
  loop.final_time = final_time;
  loop.time_step = hda::time_step_from_cfl(mf, a.data(), time_step, cfl,
                                           degree, start_time, final_time);
  const double dt = loop.time_step;
  hda::LowStorageRungeKuttaIntegrator<Number> integrator(ki, ti, rk);

  loop.loop(
    [&](double t, double step) {
      integrator.perform_time_step(op, solution, t, step);
    },
    [&](double t) {
      if (t != start_time &&
          (int)((t + 0.00000000001 - start_time) / tick) ==
            (int)((t + 0.00000000001 - start_time - dt) / tick))
        return;
      const auto e = hda::VectorTools::norm_and_error(
        mf, solution, HD_FUNC_HYPERRECTANGLE, a.data(), t);
      std::printf("   Time:%10.3e, norm: %17.10e, error: %17.10e\n", t, e[0],
                  e[1]);
    });
  return 0;
}

template <typename Number>
int dispatch(int dx, int dv, int argc, char **argv)
{
  if (dx == 1 && dv == 1)
    return run<1, 1, Number>(argc, argv);
  if (dx == 2 && dv == 2)
    return run<2, 2, Number>(argc, argv);
  if (dx == 3 && dv == 3)
    return run<3, 3, Number>(argc, argv);
  std::fprintf(stderr, "unsupported dims %d+%d\n", dx, dv);
  return 1;
}

int main(int argc, char **argv)
{
  if (argc < 14)
    {
      std::fprintf(stderr,
                   "usage: %s dim_x dim_v degree n_points collocation "
                   "periodic skew rk cells time_step final_time cfl tick "
                   "[fp32]\n",
                   argv[0]);
      return 1;
    }
  try
    {
      const int dx = std::atoi(argv[1]), dv = std::atoi(argv[2]);
      const bool fp32 = argc > 14 && std::string(argv[14]) == "fp32";
      return fp32 ? dispatch<float>(dx, dv, argc, argv)
                  : dispatch<double>(dx, dv, argc, argv);
    }
  catch (const std::exception &e)
    {
      std::fprintf(stderr, "error: %s\n", e.what());
      return 1;
    }
}
