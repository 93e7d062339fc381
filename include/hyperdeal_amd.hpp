// hyperdeal_amd.hpp -- C++ host interface over the C ABI (hyperdeal_amd.h),
// with the class names and call shapes of hyper.deal's hot path:
//
//   hyperdeal::MatrixFree                    -> hyperdeal_amd::MatrixFree
//   advection::AdvectionOperation            -> hyperdeal_amd::AdvectionOperation
//   LowStorageRungeKuttaIntegrator           -> hyperdeal_amd::LowStorageRungeKuttaIntegrator
//   TimeLoop                                 -> hyperdeal_amd::TimeLoop
//   VectorTools::interpolate/norm_and_error  -> hyperdeal_amd::VectorTools
//
// Errors raise hyperdeal_amd::Error (the reference raises
// dealii::ExceptionBase via AssertThrow).  Header only; link
// libhyperdeal_amd.so.
#ifndef HYPERDEAL_AMD_HPP
#define HYPERDEAL_AMD_HPP

#include "hyperdeal_amd.h"

#include <array>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hyperdeal_amd
{
struct Error : std::runtime_error
{
  using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what)
{
  if (rc < 0)
    throw Error(std::string(what) + ": " + hd_last_error());
}

template <typename Number>
constexpr int precision_of()
{
  return sizeof(Number) == 8 ? HD_FP64 : HD_FP32;
}

// A device vector of Number (owning).
template <typename Number>
class DeviceVector
{
public:
  DeviceVector() = default;
  explicit DeviceVector(long long n) { reinit(n); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept
    : ptr_(std::exchange(o.ptr_, nullptr)), n_(std::exchange(o.n_, 0))
  {}
  ~DeviceVector() { release(); }

  void reinit(long long n)
  {
    release();
    void *p = nullptr;
    check(hd_device_malloc(&p, (uint64_t)n * sizeof(Number)),
          "DeviceVector::reinit");
    ptr_ = static_cast<Number *>(p);
    n_ = n;
  }
  long long size() const { return n_; }
  Number *data() { return ptr_; }
  const Number *data() const { return ptr_; }

  void copy_from_host(const std::vector<Number> &v)
  {
    if ((long long)v.size() != n_)
      throw Error("DeviceVector::copy_from_host: size mismatch");
    check(hd_memcpy_htod(ptr_, v.data(), n_ * sizeof(Number)),
          "copy_from_host");
  }
  std::vector<Number> to_host() const
  {
    std::vector<Number> v(n_);
    check(hd_memcpy_dtoh(v.data(), ptr_, n_ * sizeof(Number)), "to_host");
    return v;
  }

private:
  void release()
  {
    if (ptr_)
      hd_device_free(ptr_);
    ptr_ = nullptr;
    n_ = 0;
  }
  Number *ptr_ = nullptr;
  long long n_ = 0;
};

// Phase-space matrix-free data of a Cartesian box (hyperdeal::MatrixFree,
// matrix_free/matrix_free.h:38-389).
template <int dim_x, int dim_v>
class MatrixFree
{
public:
  static constexpr int dim = dim_x + dim_v;

  struct AdditionalData
  {
    std::array<int, dim> n_cells{};
    std::array<double, dim> lo{}, hi{};
    bool periodic_x = true, periodic_v = true;
    int degree = 3;
    int n_points = 4;
    bool do_collocation = false;
  };

  MatrixFree() = default;
  MatrixFree(const MatrixFree &) = delete;
  ~MatrixFree() { hd_matrix_free_destroy(mf_); }

  void reinit(const AdditionalData &ad)
  {
    hd_mesh_desc mesh{};
    mesh.dim_x = dim_x;
    mesh.dim_v = dim_v;
    for (int d = 0; d < dim; ++d)
      {
        mesh.n_cells[d] = ad.n_cells[d];
        mesh.lo[d] = ad.lo[d];
        mesh.hi[d] = ad.hi[d];
      }
    mesh.periodic_x = ad.periodic_x;
    mesh.periodic_v = ad.periodic_v;
    hd_matrix_free_destroy(mf_);
    mf_ = nullptr;
    check(hd_matrix_free_create(&mesh, ad.degree, ad.n_points,
                                ad.do_collocation, &mf_),
          "MatrixFree::reinit");
    data_ = ad;
  }

  template <typename Number>
  void initialize_dof_vector(DeviceVector<Number> &v) const
  {
    v.reinit(n_dofs());
  }
  long long n_dofs() const { return hd_matrix_free_n_dofs(mf_); }
  long long n_cells() const { return hd_matrix_free_n_cells(mf_); }
  const AdditionalData &get_additional_data() const { return data_; }
  hd_matrix_free *handle() const { return mf_; }

private:
  hd_matrix_free *mf_ = nullptr;
  AdditionalData data_{};
};

// advection::AdvectionOperation with a constant velocity field
// (operators/advection/advection_operation.h:56-209).
template <int dim_x, int dim_v, typename Number>
class AdvectionOperation
{
public:
  static constexpr int dim = dim_x + dim_v;

  explicit AdvectionOperation(const MatrixFree<dim_x, dim_v> &data)
    : data_(data)
  {}
  AdvectionOperation(const AdvectionOperation &) = delete;
  ~AdvectionOperation() { hd_advection_destroy(op_); }

  // reinit(boundary_descriptor, velocity_field, parameters)
  void reinit(const std::array<double, dim> &velocity, double factor_skew,
              int boundary_kind = HD_BC_INHOMOGENEOUS,
              int boundary_function = HD_FUNC_ZERO)
  {
    hd_advection_destroy(op_);
    op_ = nullptr;
    velocity_ = velocity;
    check(hd_advection_create(data_.handle(), velocity.data(), factor_skew,
                              boundary_kind, boundary_function,
                              velocity.data(), &op_),
          "AdvectionOperation::reinit");
  }

  void apply(DeviceVector<Number> &dst, const DeviceVector<Number> &src,
             const double time, void *stream = nullptr) const
  {
    check_size(src, "AdvectionOperation::apply: src");
    check_size(dst, "AdvectionOperation::apply: dst");
    check(hd_advection_apply(op_, src.data(), dst.data(),
                             precision_of<Number>(), time, stream),
          "AdvectionOperation::apply");
  }

  const std::array<double, dim> &velocity() const { return velocity_; }
  hd_advection *handle() const { return op_; }
  const MatrixFree<dim_x, dim_v> &get_matrix_free() const { return data_; }

  // a vector of the operator's size (the kernels index it up to n_dofs)
  template <typename V>
  void check_size(const DeviceVector<V> &v, const char *what) const
  {
    if (v.size() != data_.n_dofs())
      throw Error(std::string(what) + ": size " + std::to_string(v.size()) +
                  " != n_dofs " + std::to_string(data_.n_dofs()));
  }

private:
  const MatrixFree<dim_x, dim_v> &data_;
  hd_advection *op_ = nullptr;
  std::array<double, dim> velocity_{};
};

// Two-register LSRK (base/time_integrators.h:47-115).  The registers hold
// Number (double, or float for the FP32-mixed form); the solution is double.
template <typename Number>
class LowStorageRungeKuttaIntegrator
{
public:
  LowStorageRungeKuttaIntegrator(DeviceVector<Number> &vec_Ki,
                                 DeviceVector<Number> &vec_Ti,
                                 const std::string &type)
    : ki_(vec_Ki), ti_(vec_Ti), type_(type)
  {
    std::array<double, 16> a{}, b{};
    n_stages_ = hd_lsrk_coefficients(type.c_str(), a.data(), b.data());
    check(n_stages_, "LowStorageRungeKuttaIntegrator");
  }

  template <typename Operator>
  void perform_time_step(const Operator &op, DeviceVector<double> &solution,
                         const double current_time, const double time_step,
                         void *stream = nullptr)
  {
    op.check_size(solution, "perform_time_step: solution");
    op.check_size(ti_, "perform_time_step: Ti");
    op.check_size(ki_, "perform_time_step: Ki");
    check(hd_lsrk_step(op.handle(), type_.c_str(), precision_of<Number>(),
                       solution.data(), ti_.data(), ki_.data(), current_time,
                       time_step, stream),
          "perform_time_step");
  }

  unsigned int n_stages() const { return n_stages_; }

private:
  DeviceVector<Number> &ki_, &ti_;
  std::string type_;
  int n_stages_ = 0;
};

// Fixed-step time loop (source/base/time_loop.cc:33-67).
struct TimeLoop
{
  double time_step = 0.1, start_time = 0.0, final_time = 20.0;
  unsigned int max_time_step_number = 100000000;

  int loop(const std::function<void(double t, double dt)> &step,
           const std::function<void(double t)> &diagnostics) const
  {
    unsigned int n = 1;
    diagnostics(start_time);
    for (double t = start_time + time_step;
         t <= final_time * 1.0000000000001 && n <= max_time_step_number;
         t += time_step, ++n)
      {
        step(t - time_step, time_step);
        diagnostics(t);
      }
    return (int)n - 1;
  }
};

struct VectorTools
{
  template <int dx, int dv>
  static void interpolate(const MatrixFree<dx, dv> &mf, int function,
                          const double *params, double time,
                          DeviceVector<double> &dst, void *stream = nullptr)
  {
    if (dst.size() != mf.n_dofs())
      throw Error("VectorTools::interpolate: size mismatch");
    check(hd_interpolate(mf.handle(), function, params, time, dst.data(),
                         stream),
          "VectorTools::interpolate");
  }

  template <int dx, int dv>
  static std::array<double, 2>
  norm_and_error(const MatrixFree<dx, dv> &mf, const DeviceVector<double> &src,
                 int function, const double *params, double time,
                 void *stream = nullptr)
  {
    if (src.size() != mf.n_dofs())
      throw Error("VectorTools::norm_and_error: size mismatch");
    std::array<double, 2> out{};
    check(hd_norm_and_error(mf.handle(), src.data(), function, params, time,
                            out.data(), stream),
          "VectorTools::norm_and_error");
    return out;
  }
};

// dt = min(TimeStep, CFL dt_crit / k^1.5), then rounded so that the last
// step ends at the final time (examples/advection/include/application.h:
// 363-396).
template <int dx, int dv>
double time_step_from_cfl(const MatrixFree<dx, dv> &mf,
                          const double *velocity, double max_time_step,
                          double cfl, int degree, double start_time,
                          double final_time)
{
  const double crit = hd_critical_time_step(mf.handle(), velocity);
  const double dt =
    std::min(max_time_step, cfl * crit / std::pow((double)degree, 1.5));
  return (final_time - start_time) /
         std::ceil((final_time - start_time) / dt);
}
} // namespace hyperdeal_amd

#endif
