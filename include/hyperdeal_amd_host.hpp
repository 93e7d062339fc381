// hyperdeal_amd_host.hpp -- the host side of hyper.deal's loop and evaluator
// interface, for operators written as cell / face / boundary lambdas.
//
// hyper.deal runs user operators through MatrixFree::loop,
// loop_cell_centric and cell_loop with std::function (or member-pointer)
// callbacks (include/hyper.deal/matrix_free/matrix_free.h:124-228,
// matrix_free.templates.h:1417-1746), inside which the phase-space
// evaluators FEEvaluation, FEFaceEvaluation and FEEvaluationInverse read
// DoFs, hold per-point data and submit integrands
// (matrix_free/fe_evaluation_cell.h:41-404, fe_evaluation_face.h:41-568,
// fe_evaluation_cell_inverse.h:36-134); the 1D sum-factorisation kernels come
// from deal.II's EvaluatorTensorProduct.  This header restates that
// interface with the same names, callback signatures and argument meanings
// for Cartesian phase-space meshes (affine cells, periodic or Dirichlet
// directions), running the callbacks on the host.  The GPU runs the built-in
// operators (AdvectionOperation, the Vlasov-Poisson operator) through the C
// ABI; device code cannot call host lambdas, so operators written against
// this interface run where it runs.
//
// Conventions (those of the reference):
//   - a cell's DoFs are its (degree+1)^D values at the Gauss-Lobatto nodes,
//     lexicographic with direction 0 fastest, x directions before v; cells
//     are numbered x first (cell = x + n_x_cells v);
//   - evaluators hold one buffer of n^D values (n = degree + 1 = n_points);
//     TensorProduct::values / gradients / hessians<dir, contract_over_rows,
//     add> apply the 1D shape values S (nodes -> Gauss points), the
//     collocation derivative at the Gauss points, and S^-1 (the inverse
//     mass matrix's factor) along one direction, contract_over_rows = true
//     in the forward sense (nodes -> points, or the derivative at points);
//   - get_gradient_x / _v read the reference-cell derivatives of direction e
//     at grad_in[q + e n_q^D]; submit_gradient_x / _v write them there,
//     multiplied by JxW and the inverse Jacobian's transpose;
//   - FEFaceEvaluation::submit_value stores -value JxW_face; the normal is
//     the outward normal of the evaluator's own cell; a face buffer holds
//     n^(D-1) values, the face's directions in increasing order.
// Only degree + 1 == n_points (Gauss points) is supported.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace hyperdeal_amd
{
namespace host
{
// ---------------------------------------------------------------------------
// 1D shape data
// ---------------------------------------------------------------------------
struct Shape1D
{
  int n = 0;                   // nodes = Gauss points = degree + 1
  std::vector<double> nodes;   // Gauss-Lobatto nodes on [0, 1]
  std::vector<double> points;  // Gauss points on [0, 1]
  std::vector<double> weights; // Gauss weights on [0, 1]
  // [dof][q] = l_dof(x_q): values<dir, true> maps nodal values to points
  std::vector<double> values;
  // [q][q'] = derivative at x_q' of the Lagrange polynomial on the Gauss
  // points that is 1 at x_q: gradients<dir, true> differentiates point values
  std::vector<double> gradients;
  // [dof][q] = (S^-1)[dof][q] with S[q][dof] = l_dof(x_q): hessians<dir,
  // true> applies S^-T, hessians<dir, false> S^-1
  std::vector<double> inverse;
};

namespace detail
{
constexpr long double kPi = 3.141592653589793238462643383279502884L;

// Legendre P_n and P_n' at z in [-1, 1]
inline void legendre(int n, long double z, long double &p, long double &dp)
{
  long double p0 = 1.0L, p1 = z;
  if (n == 0)
    {
      p = 1.0L;
      dp = 0.0L;
      return;
    }
  for (int k = 2; k <= n; ++k)
    {
      const long double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
  p = p1;
  dp = n * (z * p1 - p0) / (z * z - 1.0L);
}

inline long double lagrange(const std::vector<double> &xs, int j,
                            long double x)
{
  long double r = 1.0L;
  for (size_t m = 0; m < xs.size(); ++m)
    if ((int)m != j)
      r *= (x - xs[m]) / ((long double)xs[j] - xs[m]);
  return r;
}

inline long double lagrange_derivative(const std::vector<double> &xs, int j,
                                       long double x)
{
  long double s = 0.0L;
  for (size_t k = 0; k < xs.size(); ++k)
    {
      if ((int)k == j)
        continue;
      long double t = 1.0L / ((long double)xs[j] - xs[k]);
      for (size_t m = 0; m < xs.size(); ++m)
        if ((int)m != j && m != k)
          t *= (x - xs[m]) / ((long double)xs[j] - xs[m]);
      s += t;
    }
  return s;
}

// n x n dense inverse (Gauss-Jordan, partial pivoting, long double)
inline std::vector<double> invert(const std::vector<double> &A, int n)
{
  std::vector<long double> a(A.begin(), A.end()), b((size_t)n * n, 0.0L);
  for (int i = 0; i < n; ++i)
    b[(size_t)i * n + i] = 1.0L;
  for (int c = 0; c < n; ++c)
    {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (std::fabs((double)a[(size_t)r * n + c]) >
            std::fabs((double)a[(size_t)piv * n + c]))
          piv = r;
      for (int k = 0; k < n; ++k)
        {
          std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
          std::swap(b[(size_t)c * n + k], b[(size_t)piv * n + k]);
        }
      const long double d = a[(size_t)c * n + c];
      for (int k = 0; k < n; ++k)
        {
          a[(size_t)c * n + k] /= d;
          b[(size_t)c * n + k] /= d;
        }
      for (int r = 0; r < n; ++r)
        if (r != c)
          {
            const long double f = a[(size_t)r * n + c];
            for (int k = 0; k < n; ++k)
              {
                a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
                b[(size_t)r * n + k] -= f * b[(size_t)c * n + k];
              }
          }
    }
  return std::vector<double>(b.begin(), b.end());
}
} // namespace detail

inline Shape1D make_shape(int degree)
{
  Shape1D s;
  const int n = degree + 1;
  s.n = n;
  // Gauss points: roots of P_n (Newton from Chebyshev guesses)
  s.points.resize(n);
  s.weights.resize(n);
  for (int i = 0; i < n; ++i)
    {
      long double z = std::cos(detail::kPi * (i + 0.75L) / (n + 0.5L)), p,
                  dp;
      for (int it = 0; it < 100; ++it)
        {
          detail::legendre(n, z, p, dp);
          const long double dz = p / dp;
          z -= dz;
          if (std::fabs((double)dz) < 1e-19)
            break;
        }
      detail::legendre(n, z, p, dp);
      s.points[n - 1 - i] = (double)(0.5L * (1.0L + z));
      s.weights[n - 1 - i] = (double)(1.0L / ((1.0L - z * z) * dp * dp));
    }
  // Gauss-Lobatto nodes: 0, 1 and the roots of P_{n-1}'
  s.nodes.assign(n, 0.0);
  s.nodes[n - 1] = 1.0;
  for (int i = 1; i < n - 1; ++i)
    {
      long double z = -std::cos(detail::kPi * i / (n - 1));
      for (int it = 0; it < 100; ++it)
        {
          // f = P_{n-1}'(z), f' = P_{n-1}''(z) from the Legendre ODE
          long double p, dp;
          detail::legendre(n - 1, z, p, dp);
          const long double d2 =
            (2.0L * z * dp - (n - 1) * n * p) / (1.0L - z * z);
          const long double dz = dp / d2;
          z -= dz;
          if (std::fabs((double)dz) < 1e-19)
            break;
        }
      s.nodes[i] = (double)(0.5L * (1.0L + z));
    }
  s.values.resize((size_t)n * n);
  s.gradients.resize((size_t)n * n);
  std::vector<double> S((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int q = 0; q < n; ++q)
      {
        s.values[(size_t)j * n + q] =
          (double)detail::lagrange(s.nodes, j, s.points[q]);
        S[(size_t)q * n + j] = s.values[(size_t)j * n + q];
        s.gradients[(size_t)j * n + q] =
          (double)detail::lagrange_derivative(s.points, j, s.points[q]);
      }
  s.inverse = detail::invert(S, n);
  return s;
}

// ---------------------------------------------------------------------------
// sum factorisation along one direction (EvaluatorTensorProduct's role)
// ---------------------------------------------------------------------------
class TensorProduct
{
public:
  // dim directions of extent n; the three n x n matrices (rows: the nodal
  // or basis side, columns: the point side), any of them may be null
  TensorProduct(int dim, int n, const std::vector<double> *values,
                const std::vector<double> *gradients,
                const std::vector<double> *hessians)
    : dim_(dim), n_(n), values_(values), gradients_(gradients),
      hessians_(hessians)
  {
  }

  template <int dir, bool contract_over_rows, bool add>
  void values(const double *in, double *out) const
  {
    apply(values_, dir, contract_over_rows, add, in, out);
  }
  template <int dir, bool contract_over_rows, bool add>
  void gradients(const double *in, double *out) const
  {
    apply(gradients_, dir, contract_over_rows, add, in, out);
  }
  template <int dir, bool contract_over_rows, bool add>
  void hessians(const double *in, double *out) const
  {
    apply(hessians_, dir, contract_over_rows, add, in, out);
  }

  // the same with the direction at run time
  void apply(const std::vector<double> *M, int dir, bool contract_over_rows,
             bool add, const double *in, double *out) const
  {
    if (!M || dir < 0 || dir >= dim_)
      throw std::invalid_argument("TensorProduct: no matrix / bad direction");
    const int n = n_;
    long long stride = 1, total = 1;
    for (int d = 0; d < dim_; ++d)
      {
        if (d < dir)
          stride *= n;
        total *= n;
      }
    const long long outer = total / (stride * n);
    std::vector<double> line(n), res(n);
    for (long long o = 0; o < outer; ++o)
      for (long long s = 0; s < stride; ++s)
        {
          const long long base = o * stride * n + s;
          for (int i = 0; i < n; ++i)
            line[i] = in[base + i * stride];
          for (int i = 0; i < n; ++i)
            {
              double a = 0.0;
              for (int j = 0; j < n; ++j)
                a += contract_over_rows ? (*M)[(size_t)j * n + i] * line[j]
                                        : (*M)[(size_t)i * n + j] * line[j];
              res[i] = a;
            }
          for (int i = 0; i < n; ++i)
            out[base + i * stride] = add ? out[base + i * stride] + res[i]
                                         : res[i];
        }
  }

private:
  int dim_, n_;
  const std::vector<double> *values_, *gradients_, *hessians_;
};

// ---------------------------------------------------------------------------
// IDs and the matrix-free loops
// ---------------------------------------------------------------------------
struct TensorID
{
  // a cell (XV), a face in x (X: x = face index, v = cell) or in v (V: x =
  // cell, v = face index); inner faces come first, boundary faces after
  enum class SpaceType
  {
    XV,
    X,
    V
  };
  unsigned int x = 0, v = 0;
  unsigned int id = 0; // running number within the loop
  SpaceType type = SpaceType::XV;
};

template <int dim_x, int dim_v>
class MatrixFree
{
public:
  static constexpr int dim = dim_x + dim_v;
  using ID = TensorID;
  using SpaceType = TensorID::SpaceType;

  // a face between the cells minus and plus of one space (x or v) across
  // direction dir of that space; boundary faces have plus = -1 and side 0
  // (at the lower end of the box) or 1 (upper end)
  struct Face
  {
    int dir;
    long long minus, plus;
    int side;
  };

  // n_cells, lo, hi per phase-space direction (x first); periodic or
  // Dirichlet per space (the reference's two flags)
  void reinit(const std::array<int, dim> &n_cells,
              const std::array<double, dim> &lo,
              const std::array<double, dim> &hi, int degree,
              bool periodic_x = true, bool periodic_v = true)
  {
    for (int d = 0; d < dim; ++d)
      if (n_cells[d] < 1 || !(hi[d] > lo[d]))
        throw std::invalid_argument("MatrixFree::reinit: invalid mesh");
    n_ = n_cells;
    lo_ = lo;
    for (int d = 0; d < dim; ++d)
      h_[d] = (hi[d] - lo[d]) / n_cells[d];
    periodic_x_ = periodic_x;
    periodic_v_ = periodic_v;
    degree_ = degree;
    shape_ = make_shape(degree);
    n_x_cells_ = 1;
    n_v_cells_ = 1;
    for (int d = 0; d < dim_x; ++d)
      n_x_cells_ *= n_[d];
    for (int d = dim_x; d < dim; ++d)
      n_v_cells_ *= n_[d];
    build_faces(0, dim_x, periodic_x, faces_x_, n_inner_x_);
    build_faces(dim_x, dim_v, periodic_v, faces_v_, n_inner_v_);
  }

  int get_degree() const { return degree_; }
  int n_points() const { return shape_.n; }
  long long n_cells_x() const { return n_x_cells_; }
  long long n_cells_v() const { return n_v_cells_; }
  long long n_cells() const { return n_x_cells_ * n_v_cells_; }
  long long dofs_per_cell() const
  {
    long long r = 1;
    for (int d = 0; d < dim; ++d)
      r *= shape_.n;
    return r;
  }
  std::size_t n_dofs() const { return (size_t)(n_cells() * dofs_per_cell()); }
  const Shape1D &get_shape_info() const { return shape_; }
  double h(int d) const { return h_[d]; }
  double lo(int d) const { return lo_[d]; }
  int n_cells_dir(int d) const { return n_[d]; }
  const std::vector<Face> &faces_x() const { return faces_x_; }
  const std::vector<Face> &faces_v() const { return faces_v_; }
  long long n_inner_faces_x() const { return n_inner_x_; }
  long long n_inner_faces_v() const { return n_inner_v_; }

  void initialize_dof_vector(std::vector<double> &v) const
  {
    v.assign(n_dofs(), 0.0);
  }

  // the index per direction of cell `c` within its space (x: directions
  // 0..dim_x-1 of the x cell, v: of the v cell)
  void cell_index(long long x_cell, long long v_cell,
                  std::array<int, dim> &idx) const
  {
    for (int d = 0; d < dim_x; ++d)
      {
        idx[d] = (int)(x_cell % n_[d]);
        x_cell /= n_[d];
      }
    for (int d = dim_x; d < dim; ++d)
      {
        idx[d] = (int)(v_cell % n_[d]);
        v_cell /= n_[d];
      }
  }

  template <typename Out, typename In>
  using Op =
    std::function<void(const MatrixFree &, Out &, const In &, const ID)>;

  // every cell, x cells fastest (matrix_free.templates.h:1417-1440)
  template <typename Out, typename In>
  void cell_loop(const Op<Out, In> &cell_operation, Out &dst,
                 const In &src) const
  {
    unsigned int i0 = 0;
    for (long long v = 0; v < n_v_cells_; ++v)
      for (long long x = 0; x < n_x_cells_; ++x)
        cell_operation(*this, dst, src,
                       ID{(unsigned)x, (unsigned)v, i0++, SpaceType::XV});
  }

  // element-centric: the cell operation visits its own faces
  // (matrix_free.templates.h:1500-1600)
  template <typename Out, typename In>
  void loop_cell_centric(const Op<Out, In> &cell_operation, Out &dst,
                         const In &src) const
  {
    cell_loop(cell_operation, dst, src);
  }

  // face-centric: cells, then inner faces in x and in v, then boundary
  // faces (matrix_free.templates.h:1640-1746)
  template <typename Out, typename In>
  void loop(const Op<Out, In> &cell_operation,
            const Op<Out, In> &face_operation,
            const Op<Out, In> &boundary_operation, Out &dst,
            const In &src) const
  {
    cell_loop(cell_operation, dst, src);
    unsigned int i1 = 0, i2 = 0;
    for (long long v = 0; v < n_v_cells_; ++v)
      for (long long f = 0; f < n_inner_x_; ++f)
        face_operation(*this, dst, src,
                       ID{(unsigned)f, (unsigned)v, i1++, SpaceType::X});
    for (long long f = 0; f < n_inner_v_; ++f)
      for (long long x = 0; x < n_x_cells_; ++x)
        face_operation(*this, dst, src,
                       ID{(unsigned)x, (unsigned)f, i1++, SpaceType::V});
    for (long long v = 0; v < n_v_cells_; ++v)
      for (long long f = n_inner_x_; f < (long long)faces_x_.size(); ++f)
        boundary_operation(*this, dst, src,
                           ID{(unsigned)f, (unsigned)v, i2++, SpaceType::X});
    for (long long f = n_inner_v_; f < (long long)faces_v_.size(); ++f)
      for (long long x = 0; x < n_x_cells_; ++x)
        boundary_operation(*this, dst, src,
                           ID{(unsigned)x, (unsigned)f, i2++, SpaceType::V});
  }

  // the member-pointer forms (matrix_free.h:138-228)
  template <typename CLASS, typename Out, typename In>
  void cell_loop(void (CLASS::*cell_operation)(const MatrixFree &, Out &,
                                               const In &, const ID),
                 CLASS *owner, Out &dst, const In &src) const
  {
    cell_loop<Out, In>(bind(cell_operation, owner), dst, src);
  }
  template <typename CLASS, typename Out, typename In>
  void loop_cell_centric(void (CLASS::*cell_operation)(const MatrixFree &,
                                                       Out &, const In &,
                                                       const ID),
                         CLASS *owner, Out &dst, const In &src) const
  {
    loop_cell_centric<Out, In>(bind(cell_operation, owner), dst, src);
  }
  template <typename CLASS, typename Out, typename In>
  void loop(void (CLASS::*cell_operation)(const MatrixFree &, Out &,
                                          const In &, const ID),
            void (CLASS::*face_operation)(const MatrixFree &, Out &,
                                          const In &, const ID),
            void (CLASS::*boundary_operation)(const MatrixFree &, Out &,
                                              const In &, const ID),
            CLASS *owner, Out &dst, const In &src) const
  {
    loop<Out, In>(bind(cell_operation, owner), bind(face_operation, owner),
                  bind(boundary_operation, owner), dst, src);
  }

private:
  template <typename CLASS, typename Out, typename In>
  static Op<Out, In> bind(void (CLASS::*f)(const MatrixFree &, Out &,
                                           const In &, const ID),
                          CLASS *owner)
  {
    return [f, owner](const MatrixFree &mf, Out &dst, const In &src,
                      const ID id) { (owner->*f)(mf, dst, src, id); };
  }

  // faces of the space whose directions are [d0, d0 + nd): inner faces
  // (each once, minus = the lower cell) then boundary faces
  void build_faces(int d0, int nd, bool periodic, std::vector<Face> &faces,
                   long long &n_inner)
  {
    faces.clear();
    long long n_space = 1;
    for (int d = d0; d < d0 + nd; ++d)
      n_space *= n_[d];
    std::vector<Face> boundary;
    for (int e = 0; e < nd; ++e)
      {
        long long stride = 1;
        for (int d = d0; d < d0 + e; ++d)
          stride *= n_[d];
        const int ne = n_[d0 + e];
        for (long long c = 0; c < n_space; ++c)
          {
            const int i = (int)((c / stride) % ne);
            if (i + 1 < ne)
              faces.push_back({e, c, c + stride, -1});
            else if (periodic)
              faces.push_back({e, c, c - (long long)(ne - 1) * stride, -1});
            else
              boundary.push_back({e, c, -1, 1});
            if (!periodic && i == 0)
              boundary.push_back({e, c, -1, 0});
          }
      }
    n_inner = (long long)faces.size();
    faces.insert(faces.end(), boundary.begin(), boundary.end());
  }

  std::array<int, dim> n_{};
  std::array<double, dim> lo_{}, h_{};
  bool periodic_x_ = true, periodic_v_ = true;
  int degree_ = 0;
  Shape1D shape_;
  long long n_x_cells_ = 0, n_v_cells_ = 0;
  std::vector<Face> faces_x_, faces_v_;
  long long n_inner_x_ = 0, n_inner_v_ = 0;
};

// ---------------------------------------------------------------------------
// cell evaluators
// ---------------------------------------------------------------------------
template <int dim_x, int dim_v>
class FEEvaluation
{
public:
  static constexpr int dim = dim_x + dim_v;
  using MF = MatrixFree<dim_x, dim_v>;

  explicit FEEvaluation(const MF &mf)
    : mf_(mf), n_(mf.n_points()), nq_((int)mf.dofs_per_cell()),
      nqx_(ipow(mf.n_points(), dim_x)), data_((size_t)nq_ * (dim + 1))
  {
  }

  void reinit(const TensorID cell)
  {
    x_ = cell.x;
    v_ = cell.v;
    cell_ = x_ + mf_.n_cells_x() * v_;
    mf_.cell_index(x_, v_, idx_);
  }

  void read_dof_values(const std::vector<double> &src, double *data) const
  {
    std::copy(src.begin() + cell_ * nq_, src.begin() + (cell_ + 1) * nq_,
              data);
  }
  void read_dof_values(const std::vector<double> &src)
  {
    read_dof_values(src, data_.data());
  }
  void set_dof_values(std::vector<double> &dst) const
  {
    std::copy(data_.begin(), data_.begin() + nq_, dst.begin() + cell_ * nq_);
  }
  void distribute_local_to_global(std::vector<double> &dst) const
  {
    for (int i = 0; i < nq_; ++i)
      dst[cell_ * nq_ + i] += data_[i];
  }

  // the cell's values (first n^D entries) and room for dim more (gradients)
  double *get_data_ptr() { return data_.data(); }
  const std::vector<double> *get_shape_values() const
  {
    return &mf_.get_shape_info().values;
  }
  const std::vector<double> *get_shape_gradients() const
  {
    return &mf_.get_shape_info().gradients;
  }
  const std::vector<double> *get_inverse_shape() const
  {
    return &mf_.get_shape_info().inverse;
  }
  unsigned int n_q_points() const { return (unsigned)nq_; }
  unsigned int n_q_points_x() const { return (unsigned)nqx_; }

  double JxW(unsigned int qx, unsigned int qv) const
  {
    return jxw_space(qx, 0, dim_x) * jxw_space(qv, dim_x, dim_v);
  }

  // the x (v) gradient at point q from reference derivatives at
  // grad_in[q + e n_q^D] (fe_evaluation_cell.h:219-271)
  std::array<double, dim_x> get_gradient_x(const double *grad_in,
                                           unsigned int q, unsigned int qx,
                                           unsigned int qv) const
  {
    (void)qx;
    (void)qv;
    std::array<double, dim_x> g{};
    for (int d = 0; d < dim_x; ++d)
      g[d] = grad_in[q + (size_t)d * nq_] / mf_.h(d);
    return g;
  }
  std::array<double, dim_v> get_gradient_v(const double *grad_in,
                                           unsigned int q, unsigned int qx,
                                           unsigned int qv) const
  {
    (void)qx;
    (void)qv;
    std::array<double, dim_v> g{};
    for (int d = 0; d < dim_v; ++d)
      g[d] = grad_in[q + (size_t)d * nq_] / mf_.h(dim_x + d);
    return g;
  }

  // data[q] (+)= value JxW (fe_evaluation_cell.h:278-294)
  template <bool do_add>
  void submit_value(double *data, double value, unsigned int q,
                    unsigned int qx, unsigned int qv) const
  {
    const double v = value * JxW(qx, qv);
    data[q] = do_add ? data[q] + v : v;
  }
  // data[q + e n_q^D] = (J^-T grad)_e JxW (fe_evaluation_cell.h:301-348)
  void submit_gradient_x(double *data, const std::array<double, dim_x> &grad,
                         unsigned int q, unsigned int qx,
                         unsigned int qv) const
  {
    const double w = JxW(qx, qv);
    for (int d = 0; d < dim_x; ++d)
      data[q + (size_t)d * nq_] = grad[d] / mf_.h(d) * w;
  }
  void submit_gradient_v(double *data, const std::array<double, dim_v> &grad,
                         unsigned int q, unsigned int qx,
                         unsigned int qv) const
  {
    const double w = JxW(qx, qv);
    for (int d = 0; d < dim_v; ++d)
      data[q + (size_t)d * nq_] = grad[d] / mf_.h(dim_x + d) * w;
  }

  // (x, v) of point q (fe_evaluation_cell.h:357-397)
  std::array<double, dim> get_quadrature_point(unsigned int q) const
  {
    std::array<double, dim> p{};
    const Shape1D &s = mf_.get_shape_info();
    for (int d = 0; d < dim; ++d)
      {
        p[d] = mf_.lo(d) + (idx_[d] + s.points[q % n_]) * mf_.h(d);
        q /= n_;
      }
    return p;
  }

protected:
  static int ipow(int b, int e)
  {
    int r = 1;
    while (e-- > 0)
      r *= b;
    return r;
  }
  // product of w h over the nd directions from d0, point index q of them
  double jxw_space(unsigned int q, int d0, int nd) const
  {
    const Shape1D &s = mf_.get_shape_info();
    double w = 1.0;
    for (int d = d0; d < d0 + nd; ++d)
      {
        w *= s.weights[q % n_] * mf_.h(d);
        q /= n_;
      }
    return w;
  }

  const MF &mf_;
  int n_, nq_, nqx_;
  std::vector<double> data_;
  long long x_ = 0, v_ = 0, cell_ = 0;
  std::array<int, dim> idx_{};
};

// the inverse mass matrix's point step (fe_evaluation_cell_inverse.h:
// 123-134): data[q] /= JxW
template <int dim_x, int dim_v>
class FEEvaluationInverse : public FEEvaluation<dim_x, dim_v>
{
public:
  using FEEvaluation<dim_x, dim_v>::FEEvaluation;
  void submit_inv(double *data, unsigned int q, unsigned int qx,
                  unsigned int qv) const
  {
    data[q] /= this->JxW(qx, qv);
  }
};

// ---------------------------------------------------------------------------
// face evaluators (fe_evaluation_face.h:41-568)
// ---------------------------------------------------------------------------
template <int dim_x, int dim_v>
class FEFaceEvaluation
{
public:
  static constexpr int dim = dim_x + dim_v;
  using MF = MatrixFree<dim_x, dim_v>;
  using SpaceType = TensorID::SpaceType;

  FEFaceEvaluation(const MF &mf, bool is_minus)
    : mf_(mf), is_minus_(is_minus), n_(mf.n_points()),
      nf_((int)(mf.dofs_per_cell() / mf.n_points())), data_((size_t)nf_)
  {
  }

  // face-centric: the face of the loop (inner or boundary); this evaluator
  // takes the minus or the plus cell
  void reinit(const TensorID face)
  {
    type_ = face.type;
    if (type_ == SpaceType::XV)
      throw std::invalid_argument("FEFaceEvaluation::reinit: not a face");
    const bool in_x = type_ == SpaceType::X;
    const auto &f = in_x ? mf_.faces_x()[face.x] : mf_.faces_v()[face.v];
    boundary_ = f.plus < 0;
    if (boundary_ && !is_minus_)
      throw std::invalid_argument("FEFaceEvaluation: a boundary face has "
                                  "no plus side");
    dir_ = (in_x ? 0 : dim_x) + f.dir;
    const long long c = is_minus_ ? f.minus : f.plus;
    // the face is the upper end of the minus cell and the lower end of the
    // plus cell; a boundary face is the end given by its side
    upper_ = boundary_ ? f.side == 1 : is_minus_;
    set_cell(in_x ? c : face.x, in_x ? face.v : c);
  }

  // element-centric: face face_no = 2 d + side (side 1: the upper end) of
  // the cell; the minus evaluator takes the cell, the plus one the
  // neighbour across the face (a boundary face has none: check
  // at_boundary() of the minus evaluator first, the plus one throws)
  void reinit(const TensorID cell, unsigned int face_no)
  {
    dir_ = (int)face_no / 2;
    type_ = dir_ < dim_x ? SpaceType::X : SpaceType::V;
    const bool up = face_no % 2 == 1;
    std::array<int, dim> idx{};
    mf_.cell_index(cell.x, cell.v, idx);
    const int nd = mf_.n_cells_dir(dir_);
    boundary_ = false;
    long long x = cell.x, v = cell.v;
    if (!is_minus_)
      {
        int i = idx[dir_] + (up ? 1 : -1);
        if (i < 0 || i >= nd)
          {
            if (!periodic_dir(dir_))
              throw std::invalid_argument("FEFaceEvaluation: no neighbour "
                                          "across a boundary face");
            i = (i + nd) % nd;
          }
        idx[dir_] = i;
        x = 0;
        v = 0;
        for (int d = dim_x - 1; d >= 0; --d)
          x = x * mf_.n_cells_dir(d) + idx[d];
        for (int d = dim - 1; d >= dim_x; --d)
          v = v * mf_.n_cells_dir(d) + idx[d];
      }
    else
      boundary_ = !periodic_dir(dir_) &&
                  ((up && idx[dir_] == nd - 1) || (!up && idx[dir_] == 0));
    upper_ = is_minus_ ? up : !up;
    set_cell(x, v);
  }

  bool at_boundary() const { return boundary_; }
  SpaceType get_type() const { return type_; }
  double *get_data_ptr() { return data_.data(); }
  const std::vector<double> *get_shape_values() const
  {
    return &mf_.get_shape_info().values;
  }

  // the n^(D-1) nodal values of the cell on the face
  void read_dof_values(const std::vector<double> &src)
  {
    read_dof_values_from_buffer(src.data() + cell_ * mf_.dofs_per_cell());
  }
  void read_dof_values_from_buffer(const double *cell_data)
  {
    for (int i = 0; i < nf_; ++i)
      data_[i] = cell_data[cell_node(i)];
  }
  void distribute_to_buffer(double *cell_data) const
  {
    for (int i = 0; i < nf_; ++i)
      cell_data[cell_node(i)] += data_[i];
  }
  void distribute_local_to_global(std::vector<double> &dst) const
  {
    distribute_to_buffer(dst.data() + cell_ * mf_.dofs_per_cell());
  }

  // data[q] = -value JxW_face (fe_evaluation_face.h:381-399); q = qx + (x
  // face points) qv
  template <SpaceType stype>
  void submit_value(double *data, double value, unsigned int q,
                    unsigned int qx, unsigned int qv) const
  {
    if (stype != type_)
      throw std::invalid_argument("FEFaceEvaluation: face types differ");
    data[q] = -value * jxw(qx, qv);
  }

  // the outward normal of this evaluator's cell (fe_evaluation_face.h:
  // 434-463): zero in the space the face does not cut
  std::array<double, dim_x> get_normal_vector_x(unsigned int qx) const
  {
    (void)qx;
    std::array<double, dim_x> n{};
    if (dir_ < dim_x)
      n[dir_] = upper_ ? 1.0 : -1.0;
    return n;
  }
  std::array<double, dim_v> get_normal_vector_v(unsigned int qv) const
  {
    (void)qv;
    std::array<double, dim_v> n{};
    if (dir_ >= dim_x)
      n[dir_ - dim_x] = upper_ ? 1.0 : -1.0;
    return n;
  }

  // (x, v) of face point q
  template <SpaceType stype>
  std::array<double, dim> get_quadrature_point(unsigned int qx,
                                               unsigned int qv) const
  {
    if (stype != type_)
      throw std::invalid_argument("FEFaceEvaluation: face types differ");
    const unsigned int nqx = face_points_x();
    unsigned int q = qx + nqx * qv;
    const Shape1D &s = mf_.get_shape_info();
    std::array<double, dim> p{};
    for (int d = 0; d < dim; ++d)
      {
        double t;
        if (d == dir_)
          t = upper_ ? 1.0 : 0.0;
        else
          {
            t = s.points[q % n_];
            q /= n_;
          }
        p[d] = mf_.lo(d) + (idx_[d] + t) * mf_.h(d);
      }
    return p;
  }

  // face points in x (the q = qx + face_points_x() qv split)
  unsigned int face_points_x() const
  {
    int r = 1;
    for (int d = 0; d < dim_x - (type_ == SpaceType::X ? 1 : 0); ++d)
      r *= n_;
    return (unsigned)r;
  }

private:
  bool periodic_dir(int d) const
  {
    // a direction is periodic when its space has no boundary faces
    const bool in_x = d < dim_x;
    const auto &faces = in_x ? mf_.faces_x() : mf_.faces_v();
    const long long inner = in_x ? mf_.n_inner_faces_x() : mf_.n_inner_faces_v();
    return (long long)faces.size() == inner;
  }

  void set_cell(long long x, long long v)
  {
    cell_ = x + mf_.n_cells_x() * v;
    mf_.cell_index(x, v, idx_);
  }

  // cell-local node of face entry i (the face's directions in increasing
  // order, the normal direction at its end)
  int cell_node(int i) const
  {
    int node = 0, stride = 1;
    for (int d = 0; d < dim; ++d)
      {
        int k;
        if (d == dir_)
          k = upper_ ? n_ - 1 : 0;
        else
          {
            k = i % n_;
            i /= n_;
          }
        node += k * stride;
        stride *= n_;
      }
    return node;
  }

  // JxW of face point (qx, qv): w h over the face's directions
  double jxw(unsigned int qx, unsigned int qv) const
  {
    unsigned int q = qx + face_points_x() * qv;
    const Shape1D &s = mf_.get_shape_info();
    double w = 1.0;
    for (int d = 0; d < dim; ++d)
      if (d != dir_)
        {
          w *= s.weights[q % n_] * mf_.h(d);
          q /= n_;
        }
    return w;
  }

  const MF &mf_;
  bool is_minus_;
  int n_, nf_;
  std::vector<double> data_;
  SpaceType type_ = SpaceType::X;
  int dir_ = 0;
  bool upper_ = true, boundary_ = false;
  long long cell_ = 0;
  std::array<int, dim> idx_{};
};
} // namespace host
} // namespace hyperdeal_amd
