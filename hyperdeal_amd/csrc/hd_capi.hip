// hd_capi.hip -- the extern "C" boundary (include/hyperdeal_amd.h): handles,
// argument checks, algorithm selection and kernel dispatch.
#include "../../include/hyperdeal_amd.h"
#include "hd_device.h"
#include "hd_functions.h"
#include "hd_handles.h"
#include "hd_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace hd_abi
{
thread_local std::string g_last_error;
} // namespace hd_abi

using hd_abi::fail;
using hd_abi::hip_fail;
using hd_abi::ok;

namespace
{
using hd_abi::as_stream;

int choose_algorithm(const hd_advection *op, int precision)
{
  const hd_matrix_free *mf = op->mf;
  (void)precision; // both kernels have FP64 and FP32-mixed forms
  const bool quad_ok = hd::quadrature_supported(mf->D, mf->nd, mf->nq);
  const bool coll_ok = op->collapsed_ok;
  switch (op->algorithm)
    {
      case HD_ALGO_FCL:
        // deformed: the inverse mass matrix uses the operator's points
        return !op->vp && hd::fcl_supported(mf->D, mf->nd, mf->nq) &&
                   (!mf->deformed || mf->nq == mf->nd)
                 ? HD_ALGO_FCL
                 : -1;
      case HD_ALGO_COLLAPSED:
        return coll_ok ? HD_ALGO_COLLAPSED : -1;
      case HD_ALGO_QUADRATURE:
        return quad_ok ? HD_ALGO_QUADRATURE : -1;
      default:
        return coll_ok ? HD_ALGO_COLLAPSED : (quad_ok ? HD_ALGO_QUADRATURE : -1);
    }
}

// where a Vlasov-Poisson stage takes its density from
enum VpDensity
{
  kDensityInput = 0,    // the stage input (FP64, or FP32 in the mixed form)
  kDensitySolution = 1, // the FP64 solution (first stage of an LSRK step)
  kDensityPartials = 2, // the partials the previous stage's kernel wrote
  kDensityReady = 3     // rho is computed already (with the step's copy)
};

// the phase-space kernel runs (not the quadrature kernel)
bool vp_kernel_active(const hd_advection *op)
{
  return op->vp && op->vtab && op->algorithm != HD_ALGO_QUADRATURE;
}

// rho (from f, or from the partials of the previous stage), the potential
// solve (with the learned guess when it is on), grad phi
int vp_update(hd_advection *op, const double *f, const float *f32,
              hipStream_t stream, bool from_partials = false,
              bool rho_ready = false)
{
  hipError_t e =
    rho_ready ? hipSuccess
    : from_partials
      ? hd::vp_density_from_partials(op->rho_part, op->n_v_cells,
                                     op->n_x_dofs, op->rho, stream)
      : hd::vp_density(f, f32, op->vpdev, op->n_x_cells, op->n_v_cells,
                       op->partial, op->density_chunks, op->rho, stream);
  const double *guess = nullptr;
  if (e == hipSuccess && op->learned)
    e = hd_abi::learned_before_solve(*op, &guess, stream);
  else if (e == hipSuccess && op->fdm_on)
    {
      e = hd::vp_fdm_solve(*op->fdm, op->ptab, op->rho, stream);
      guess = op->fdm->x;
    }
  if (e == hipSuccess)
    e = hd::vp_poisson(op->xdev, op->mf->nd, op->rho, op->phi, guess,
                       op->efield, op->work, op->poisson_tol,
                       op->poisson_max_it, op->stats, stream, op->ptab,
                       nullptr, !op->learned && op->fdm_on);
  if (e == hipSuccess && op->learned)
    e = hd_abi::learned_after_solve(*op, stream);
  return e == hipSuccess ? HD_OK : hip_fail(e, "Vlasov-Poisson field update");
}

// density: where a Vlasov-Poisson operator takes rho from (VpDensity);
// sol: the FP64 solution (kDensitySolution)
int run_operator(hd_advection *op, int mode, int precision, const void *src,
                 void *dst, const hd::StageDev &stage, double time,
                 hipStream_t stream, int density = kDensityInput,
                 const double *sol = nullptr)
{
  hd_matrix_free *mf = op->mf;
  if (op->vp && op->vp_auto)
    {
      int rc;
      if (density == kDensityReady)
        rc = vp_update(op, nullptr, nullptr, stream, false, true);
      else if (density == kDensityPartials)
        rc = vp_update(op, nullptr, nullptr, stream, true);
      else if (density == kDensitySolution)
        rc = vp_update(op, sol, nullptr, stream);
      else if (precision == HD_FP64)
        rc = vp_update(op, static_cast<const double *>(src), nullptr, stream);
      else
        rc = vp_update(op, nullptr, static_cast<const float *>(src), stream);
      if (rc != HD_OK)
        return rc;
    }
  if (vp_kernel_active(op))
    {
      const hipError_t e =
        hd::vp_wave_apply(mf->nd, mode, precision == HD_FP64 ? 0 : 1, src,
                          dst, op->vtab, mf->mesh, op->n_x_cells, op->efield,
                          stage, stream);
      return e == hipSuccess ? ok() : hip_fail(e, "kernel launch");
    }
  const int algo = choose_algorithm(op, precision);
  if (algo < 0)
    return fail(HD_ERR_UNSUPPORTED,
                "no kernel for this operator (dims " + std::to_string(mf->dim_x) +
                  "+" + std::to_string(mf->dim_v) + ", degree " +
                  std::to_string(mf->degree) + ", n_points " +
                  std::to_string(mf->nq) + ", precision " +
                  std::to_string(precision) + ", algorithm " +
                  std::to_string(op->algorithm) + ")");
  hipError_t e;
  if (algo == HD_ALGO_FCL)
    {
      // the integrals are assembled in dst (apply), in ti_out (stage), or
      // in a free vector (last stage)
      void *asmv = mode == 0 ? dst : stage.ti_out;
      if (!asmv)
        asmv = stage.scratch;
      if (!asmv)
        {
          if (!op->split_scratch)
            {
              e = hipMalloc(&op->split_scratch,
                            (size_t)mf->n_dofs * sizeof(double));
              if (e != hipSuccess)
                return hip_fail(e, "FCL assembly vector");
            }
          asmv = op->split_scratch;
        }
      if (precision == HD_FP64)
        e = hd::fcl_apply(mf->D, mf->nd, mf->nq, mode,
                          static_cast<const double *>(src),
                          static_cast<double *>(asmv),
                          static_cast<double *>(dst), op->quad, op->mass,
                          mf->mesh, stage, time, mf->dim_x, mf->collocation,
                          op->geo, stream);
      else
        e = hd::fcl_apply_f32(mf->D, mf->nd, mf->nq, mode,
                              static_cast<const float *>(src),
                              static_cast<float *>(asmv),
                              static_cast<float *>(dst), op->quad, op->mass,
                              mf->mesh, stage, time, mf->dim_x,
                              mf->collocation, op->geo, stream);
    }
  else if (algo == HD_ALGO_COLLAPSED)
    {
      // the cell and face loads are 16-byte vector loads
      if (reinterpret_cast<uintptr_t>(src) % 16 != 0)
        return fail(HD_ERR_INVALID,
                    "the collapsed kernel needs a 16-byte aligned source "
                    "vector");
      if (op->split && precision != HD_FP64 &&
          hd::collapsed_cres_supported(mf->D, mf->nd))
        e = hd::collapsed_cres_f32(mf->nd, mode,
                                   static_cast<const float *>(src),
                                   static_cast<float *>(dst), op->coef32,
                                   mf->mesh, op->inflow, stage, stream);
      else if (op->split)
        {
          // the first part lands in dst (apply), in ti_out (stage), or in a
          // free vector (last stage)
          void *part = mode == 0 ? dst : stage.ti_out;
          if (!part)
            part = stage.scratch;
          if (!part)
            {
              if (!op->split_scratch)
                {
                  e = hipMalloc(&op->split_scratch,
                                (size_t)mf->n_dofs * sizeof(double));
                  if (e != hipSuccess)
                    return hip_fail(e, "split-form scratch vector");
                }
              part = op->split_scratch;
            }
          if (precision == HD_FP64)
            e = hd::collapsed_split_f64(
              mf->D, mf->nd, mode, static_cast<const double *>(src),
              static_cast<double *>(dst), static_cast<double *>(part),
              op->coef64, mf->mesh, op->inflow, stage, stream);
          else
            e = hd::collapsed_split_f32(
              mf->D, mf->nd, mode, static_cast<const float *>(src),
              static_cast<float *>(dst), static_cast<float *>(part),
              op->coef32, mf->mesh, op->inflow, stage, stream);
        }
      else if (hd::collapsed_wave_supported(mf->D, mf->nd,
                                            precision == HD_FP64 ? 0 : 1))
        e = precision == HD_FP64
              ? hd::collapsed_wave_f64(mf->nd, mode,
                                       static_cast<const double *>(src),
                                       static_cast<double *>(dst),
                                       op->coef64, mf->mesh, op->inflow,
                                       stage, stream)
              : hd::collapsed_wave_f32(mf->nd, mode,
                                       static_cast<const float *>(src),
                                       static_cast<float *>(dst), op->coef32,
                                       mf->mesh, op->inflow, stage, stream);
      else if (precision == HD_FP64)
        e = hd::collapsed_apply_f64(mf->D, mf->nd, mode,
                                    static_cast<const double *>(src),
                                    static_cast<double *>(dst), op->coef64,
                                    mf->mesh, op->inflow, stage, stream);
      else
        e = hd::collapsed_apply_f32(mf->D, mf->nd, mode,
                                    static_cast<const float *>(src),
                                    static_cast<float *>(dst), op->coef32,
                                    mf->mesh, op->inflow, stage, stream);
      // Dirichlet faces: the kernel read the periodic images there
      if (e == hipSuccess && op->dirichlet)
        e = precision == HD_FP64
              ? hd::dirichlet_correct_f64(
                  mf->D, mf->nd, mf->nq, mode, static_cast<const double *>(src),
                  static_cast<double *>(dst), op->coef64, mf->mesh, op->inflow,
                  op->quad, stage, time, mf->dim_x, stream)
              : hd::dirichlet_correct_f32(
                  mf->D, mf->nd, mf->nq, mode, static_cast<const float *>(src),
                  static_cast<float *>(dst), op->coef32, mf->mesh, op->inflow,
                  op->quad, stage, time, mf->dim_x, stream);
    }
  else if (precision == HD_FP64)
    e = hd::quadrature_apply(mf->D, mf->nd, mf->nq, mode,
                             static_cast<const double *>(src),
                             static_cast<double *>(dst), op->quad, mf->mesh,
                             stage, time, mf->dim_x, mf->collocation,
                             op->vp ? op->efield : nullptr, op->geo, stream);
  else
    e = hd::quadrature_apply_f32(mf->D, mf->nd, mf->nq, mode,
                                 static_cast<const float *>(src),
                                 static_cast<float *>(dst), op->quad,
                                 mf->mesh, stage, time, mf->dim_x,
                                 mf->collocation,
                                 op->vp ? op->efield : nullptr, op->geo,
                                 stream);
  if (e != hipSuccess)
    return hip_fail(e, "kernel launch");
  return ok();
}

bool valid_precision(int p) { return p == HD_FP64 || p == HD_FP32; }
size_t precision_bytes(int p) { return p == HD_FP64 ? 8 : 4; }

bool rk_table(const char *type, std::vector<double> &a, std::vector<double> &b)
{
  // Kennedy, Carpenter, Lewis (2000) two-register schemes, the values of
  // base/time_integrators.templates.h:32-84
  const std::string t = type ? type : "";
  if (t == "rk33")
    {
      b = {0.245170287303492, 0.184896052186740, 0.569933660509768};
      a = {0.755726351946097, 0.386954477304099};
    }
  else if (t == "rk45")
    {
      b = {1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
           -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
           5198255086312. / 14908931495163.};
      a = {970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
           2251764453980. / 15575788980749.,
           26877169314380. / 34165994151039.};
    }
  else if (t == "rk47")
    {
      b = {0.0941840925477795334, 0.149683694803496998, 0.285204742060440058,
           -0.122201846148053668, 0.0605151571191401122, 0.345986987898399296,
           0.186627171718797670};
      const double g[6] = {0.241566650129646868, 0.0423866513027719953,
                           0.215602732678803776, 0.232328007537583987,
                           0.256223412574146438, 0.0978694102142697230};
      a.assign(6, 0.0);
      for (int i = 0; i < 6; ++i)
        a[i] = g[i] + b[i];
    }
  else if (t == "rk59")
    {
      b = {2274579626619. / 23610510767302., 693987741272. / 12394497460941.,
           -347131529483. / 15096185902911., 1144057200723. / 32081666971178.,
           1562491064753. / 11797114684756., 13113619727965. / 44346030145118.,
           393957816125. / 7825732611452., 720647959663. / 6565743875477.,
           3559252274877. / 14424734981077.};
      a = {1107026461565. / 5417078080134.,  38141181049399. / 41724347789894.,
           493273079041. / 11940823631197.,  1851571280403. / 6147804934346.,
           11782306865191. / 62590030070788., 9452544825720. / 13648368537481.,
           4435885630781. / 26285702406235., 2357909744247. / 11371140753790.};
    }
  else
    return false;
  return true;
}
} // namespace

extern "C" {

const char *hd_last_error(void) { return hd_abi::g_last_error.c_str(); }

int hd_version(int *major, int *minor)
{
  if (major)
    *major = 0;
  if (minor)
    *minor = 1;
  return ok();
}

int hd_matrix_free_create(const hd_mesh_desc *mesh, int degree, int n_points,
                          int collocation, hd_matrix_free **out)
{
  if (!mesh || !out)
    return fail(HD_ERR_INVALID, "hd_matrix_free_create: null argument");
  *out = nullptr;
  if (mesh->dim_x < 1 || mesh->dim_x > 3 || mesh->dim_v < 1 || mesh->dim_v > 3)
    return fail(HD_ERR_INVALID, "dim_x and dim_v must be in 1..3");
  if (degree < 1 || degree > 7 || n_points < degree + 1 || n_points > 8 ||
      (collocation && n_points != degree + 1))
    return fail(HD_ERR_INVALID, "invalid degree / n_points / collocation");
  auto *mf = new (std::nothrow) hd_matrix_free();
  if (!mf)
    return fail(HD_ERR_NOMEM, "out of host memory");
  mf->dim_x = mesh->dim_x;
  mf->dim_v = mesh->dim_v;
  mf->D = mesh->dim_x + mesh->dim_v;
  mf->degree = degree;
  mf->nd = degree + 1;
  mf->nq = n_points;
  mf->collocation = collocation ? 1 : 0;
  if (!mf->tables.build(degree, n_points, collocation != 0))
    {
      delete mf;
      return fail(HD_ERR_INVALID, "cannot build 1D tables");
    }
  mf->mesh.dim = mf->D;
  long long nc = 1;
  for (int d = 0; d < mf->D; ++d)
    {
      if (mesh->n_cells[d] < 1 || !(mesh->hi[d] > mesh->lo[d]))
        {
          delete mf;
          return fail(HD_ERR_INVALID, "invalid box or cell count");
        }
      mf->mesh.n[d] = mesh->n_cells[d];
      mf->mesh.stride[d] = nc;
      mf->mesh.lo[d] = mesh->lo[d];
      mf->mesh.h[d] = (mesh->hi[d] - mesh->lo[d]) / mesh->n_cells[d];
      mf->periodic[d] = d < mf->dim_x ? mesh->periodic_x != 0
                                      : mesh->periodic_v != 0;
      nc *= mesh->n_cells[d];
    }
  if (nc >= (1LL << 31))
    {
      delete mf;
      return fail(HD_ERR_UNSUPPORTED, "more than 2^31 cells");
    }
  mf->mesh.n_cells = nc;
  mf->dofs_cell = hd::ipow(mf->nd, mf->D);
  mf->n_dofs = nc * mf->dofs_cell;
  *out = mf;
  return ok();
}

int hd_matrix_free_destroy(hd_matrix_free *mf)
{
  if (!mf)
    return ok();
  if (mf->n_ops > 0)
    return fail(HD_ERR_INVALID,
                "hd_matrix_free_destroy: operators of this object are alive");
  if (mf->partial)
    (void)hipFree(mf->partial);
  if (mf->geo_jw)
    (void)hipFree(mf->geo_jw);
  delete mf;
  return ok();
}

int hd_matrix_free_set_deformation(hd_matrix_free *mf, double deformation_x,
                                   int frequency_x, double deformation_v,
                                   int frequency_v)
{
  if (!mf || frequency_x < 1 || frequency_v < 1)
    return fail(HD_ERR_INVALID, "hd_matrix_free_set_deformation: invalid "
                                "argument");
  if (mf->n_ops > 0)
    return fail(HD_ERR_INVALID,
                "hd_matrix_free_set_deformation: operators already exist on "
                "this object (deform the mesh before creating them)");
  if (mf->geo_jw)
    (void)hipFree(mf->geo_jw);
  mf->geo_jw = nullptr;
  mf->geo = hd::DeformDev{};
  mf->deformed = deformation_x != 0.0 || deformation_v != 0.0;
  if (!mf->deformed)
    return ok();
  hd::SpaceMap *maps[2] = {&mf->xmap, &mf->vmap};
  const int b[2] = {0, mf->dim_x}, dims[2] = {mf->dim_x, mf->dim_v};
  const double delta[2] = {deformation_x, deformation_v};
  const int freq[2] = {frequency_x, frequency_v};
  std::vector<double> jw[2], unused;
  for (int s = 0; s < 2; ++s)
    {
      hd::SpaceMap &m = *maps[s];
      m.dim = dims[s];
      for (int j = 0; j < m.dim; ++j)
        {
          const int d = b[s] + j;
          m.n[j] = mf->mesh.n[d];
          m.lo[j] = mf->mesh.lo[d];
          m.h[j] = mf->mesh.h[d];
          m.hi[j] = mf->mesh.lo[d] + mf->mesh.n[d] * mf->mesh.h[d];
        }
      m.delta = delta[s];
      m.freq = freq[s];
      hd::space_tables(m, mf->nq, mf->tables.qpoints.data(),
                       mf->tables.qweights.data(), nullptr, jw[s], unused);
    }
  const size_t nx = jw[0].size(), nv = jw[1].size();
  hipError_t e = hipMalloc(&mf->geo_jw, (nx + nv) * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpy(mf->geo_jw, jw[0].data(), nx * sizeof(double),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(mf->geo_jw + nx, jw[1].data(), nv * sizeof(double),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess)
    {
      mf->deformed = false;
      return hip_fail(e, "hd_matrix_free_set_deformation");
    }
  hd::DeformDev &g = mf->geo;
  g.on = 1;
  g.dim_x = mf->dim_x;
  g.dim_v = mf->dim_v;
  g.delta[0] = deformation_x;
  g.delta[1] = deformation_v;
  g.freq[0] = frequency_x;
  g.freq[1] = frequency_v;
  g.n_x_cells = mf->xmap.n_cells();
  g.xjw = mf->geo_jw;
  g.vjw = mf->geo_jw + nx;
  return ok();
}

long long hd_matrix_free_n_dofs(const hd_matrix_free *mf)
{
  return mf ? mf->n_dofs : -1;
}
long long hd_matrix_free_n_cells(const hd_matrix_free *mf)
{
  return mf ? mf->mesh.n_cells : -1;
}
int hd_matrix_free_dofs_per_cell(const hd_matrix_free *mf)
{
  return mf ? mf->dofs_cell : -1;
}

int hd_advection_create(hd_matrix_free *mf, const double *velocity,
                        double skew_factor, int bc_kind, int bc_function,
                        const double *bc_params, hd_advection **out)
{
  if (!mf || !velocity || !out)
    return fail(HD_ERR_INVALID, "hd_advection_create: null argument");
  *out = nullptr;
  if (bc_kind != HD_BC_HOMOGENEOUS && bc_kind != HD_BC_INHOMOGENEOUS)
    return fail(HD_ERR_INVALID, "invalid boundary kind");
  auto *op = new (std::nothrow) hd_advection();
  if (!op)
    return fail(HD_ERR_NOMEM, "out of host memory");
  op->mf = mf;
  op->skew = skew_factor;
  op->bc_kind = bc_kind;
  op->bc_function = bc_function;
  const int D = mf->D, nd = mf->nd, nq = mf->nq;
  for (int d = 0; d < D; ++d)
    {
      op->a[d] = velocity[d];
      op->bc_params[d] = bc_params ? bc_params[d] : velocity[d];
    }

  // collapsed form: Cartesian; Dirichlet directions through the boundary
  // correction after the collapsed kernel
  bool all_periodic = true;
  for (int d = 0; d < D; ++d)
    all_periodic = all_periodic && mf->periodic[d];
  op->dirichlet = !all_periodic;
  op->split = !hd::collapsed_supported(D, nd) &&
              hd::collapsed_split_supported(D, nd);
  op->collapsed_ok =
    !mf->deformed &&
    (all_periodic ||
     hd::dirichlet_collapsed_supported(D, nq,
                                       bc_kind == HD_BC_INHOMOGENEOUS)) &&
    (hd::collapsed_supported(D, nd) || op->split);
  if (op->collapsed_ok)
    {
      const int per = nd * nd + nd;
      std::vector<double> c64(D * per);
      std::vector<float> c32(D * per);
      for (int d = 0; d < D; ++d)
        {
          const auto c = hd::collapse_direction(mf->tables, op->a[d],
                                                mf->mesh.h[d], skew_factor);
          op->inflow.side[d] =
            op->a[d] == 0.0 ? 0 : (c.inflow_from_left ? 1 : -1);
          std::copy(c.A.begin(), c.A.end(), c64.begin() + d * per);
          std::copy(c.b.begin(), c.b.end(), c64.begin() + d * per + nd * nd);
        }
      for (size_t i = 0; i < c64.size(); ++i)
        c32[i] = (float)c64[i];
      hipError_t e = hipMalloc(&op->coef64, c64.size() * sizeof(double));
      if (e == hipSuccess)
        e = hipMalloc(&op->coef32, c32.size() * sizeof(float));
      if (e == hipSuccess)
        e = hipMemcpy(op->coef64, c64.data(), c64.size() * sizeof(double),
                      hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(op->coef32, c32.data(), c32.size() * sizeof(float),
                      hipMemcpyHostToDevice);
      if (e != hipSuccess)
        {
          hd_advection_destroy(op);
          return hip_fail(e, "hd_advection_create");
        }
    }

  // quadrature form: the 1D tables by value
  if (nq <= 8)
    {
      auto &q = op->quad;
      const auto &t = mf->tables;
      std::copy(t.S.begin(), t.S.end(), q.S);
      std::copy(t.Dq.begin(), t.Dq.end(), q.Dq);
      std::copy(t.F.begin(), t.F.end(), q.F);
      std::copy(t.Sinv.begin(), t.Sinv.end(), q.Sinv);
      std::copy(t.qweights.begin(), t.qweights.end(), q.w);
      std::copy(t.qpoints.begin(), t.qpoints.end(), q.qp);
      for (int d = 0; d < D; ++d)
        {
          q.a[d] = op->a[d];
          q.periodic[d] = mf->periodic[d] ? 1 : 0;
          q.bc_params[d] = op->bc_params[d];
        }
      q.skew = skew_factor;
      q.bc_inhomogeneous = bc_kind == HD_BC_INHOMOGENEOUS ? 1 : 0;
      q.bc_function = bc_function;
      // the FCL inverse mass matrix: the Gauss(k+1) rule (the reference's
      // inverse-mass evaluator uses quadrature no. 1), or the Gauss-Lobatto
      // rule of the nodes for collocation
      auto &m = op->mass;
      m.collocation = mf->collocation;
      hd::Tables1D g;
      const hd::Tables1D *tm = &t;
      if (!mf->collocation && nq != nd)
        {
          if (!g.build(mf->degree, nd, false))
            {
              hd_advection_destroy(op);
              return fail(HD_ERR_INVALID, "cannot build the mass tables");
            }
          tm = &g;
        }
      std::copy(tm->Sinv.begin(), tm->Sinv.end(), m.Sinv);
      std::copy(tm->qweights.begin(), tm->qweights.end(), m.w);
    }

  // deformed mesh: J^{-1} a and JxW at the cell points, (a . n) JxW_face
  // at the face points, per space
  if (mf->deformed)
    {
      std::vector<double> xc, xf, vc, vf;
      const auto &t = mf->tables;
      hd::space_tables(mf->xmap, nq, t.qpoints.data(), t.qweights.data(),
                       op->a, xc, xf);
      hd::space_tables(mf->vmap, nq, t.qpoints.data(), t.qweights.data(),
                       op->a + mf->dim_x, vc, vf);
      const size_t n = xc.size() + xf.size() + vc.size() + vf.size();
      hipError_t e = hipMalloc(&op->geo_tab, n * sizeof(double));
      size_t off = 0;
      double *parts[4] = {nullptr, nullptr, nullptr, nullptr};
      const std::vector<double> *src[4] = {&xc, &xf, &vc, &vf};
      for (int i = 0; i < 4 && e == hipSuccess; ++i)
        {
          parts[i] = op->geo_tab + off;
          e = hipMemcpy(parts[i], src[i]->data(),
                        src[i]->size() * sizeof(double),
                        hipMemcpyHostToDevice);
          off += src[i]->size();
        }
      if (e != hipSuccess)
        {
          hd_advection_destroy(op);
          return hip_fail(e, "hd_advection_create: deformed-mesh tables");
        }
      op->geo = mf->geo;
      op->geo.xca = parts[0];
      op->geo.xfg = parts[1];
      op->geo.vca = parts[2];
      op->geo.vfg = parts[3];
    }
  mf->n_ops += 1;
  op->counted = true;
  *out = op;
  return ok();
}

int hd_advection_destroy(hd_advection *op)
{
  if (!op)
    return ok();
  for (void *p : {(void *)op->coef64, (void *)op->coef32, (void *)op->efield,
                  (void *)op->rho, (void *)op->phi, (void *)op->work,
                  (void *)op->partial, (void *)op->stats,
                  (void *)op->ptab, (void *)op->vtab, (void *)op->rho_part,
                  op->split_scratch, (void *)op->geo_tab})
    if (p)
      (void)hipFree(p);
  if (op->mf && op->mf->n_ops > 0 && op->counted)
    op->mf->n_ops -= 1;
  hd_abi::learned_release(op->learned);
  if (op->fdm)
    {
      hd::vp_fdm_destroy(*op->fdm);
      delete op->fdm;
    }
  delete op;
  return ok();
}

int hd_vlasov_poisson_create(hd_matrix_free *mf, double skew_factor,
                             double poisson_tol, hd_advection **out)
{
  if (!mf || !out)
    return fail(HD_ERR_INVALID, "hd_vlasov_poisson_create: null argument");
  *out = nullptr;
  if (mf->dim_x != mf->dim_v)
    return fail(HD_ERR_UNSUPPORTED,
                "Vlasov-Poisson needs dim_x == dim_v (the reference's "
                "v-integration assumes it)");
  for (int d = 0; d < mf->D; ++d)
    if (!mf->periodic[d])
      return fail(HD_ERR_UNSUPPORTED, "Vlasov-Poisson needs a periodic box");
  if (mf->deformed)
    return fail(HD_ERR_UNSUPPORTED,
                "Vlasov-Poisson runs on Cartesian meshes only");
  if (mf->nq != mf->nd || !hd::quadrature_supported(mf->D, mf->nd, mf->nq))
    return fail(HD_ERR_UNSUPPORTED,
                "Vlasov-Poisson runs the quadrature kernel: n_points must "
                "be degree + 1");
  long long ncx = 1;
  for (int d = 0; d < mf->dim_x; ++d)
    ncx *= mf->mesh.n[d];
  if (!hd::vp_poisson_supported(mf->dim_x, mf->nd, ncx))
    return fail(HD_ERR_UNSUPPORTED, "x mesh too large for the Poisson solver");
  // the operator: constant part unused, velocity from the field
  std::vector<double> zero(mf->D, 0.0);
  int rc = hd_advection_create(mf, zero.data(), skew_factor,
                               HD_BC_HOMOGENEOUS, HD_FUNC_ZERO, nullptr, out);
  if (rc != HD_OK)
    return rc;
  hd_advection *op = *out;
  op->vp = true;
  op->collapsed_ok = false;
  op->poisson_tol = poisson_tol > 0 ? poisson_tol : 1e-10;
  op->quad.vp = 1;
  op->quad.n_x_cells = ncx;
  op->n_x_cells = ncx;
  op->n_v_cells = mf->mesh.n_cells / ncx;
  const int nd = mf->nd, nq = mf->nq;
  const auto &t = mf->tables;
  op->n_x_dofs = ncx * hd::ipow(nd, mf->dim_x);
  // x-space matrices (exact: the Gauss rule has k+1 points)
  auto &x = op->xdev;
  x.dim = mf->dim_x;
  x.n_cells = ncx;
  x.kk = std::max(mf->degree, 1) * (mf->degree + 1.0);
  for (int d = 0; d < mf->dim_x; ++d)
    {
      x.n[d] = mf->mesh.n[d];
      x.h[d] = mf->mesh.h[d];
    }
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j)
      {
        double m = 0.0, k = 0.0;
        for (int q = 0; q < nq; ++q)
          {
            m += t.qweights[q] * t.S[q * nd + i] * t.S[q * nd + j];
            k += t.qweights[q] * t.G[q * nd + i] * t.G[q * nd + j];
          }
        x.M[i * nd + j] = m;
        x.K[i * nd + j] = k;
      }
  for (int i = 0; i < nd; ++i)
    {
      x.d0[i] = t.dnode0[i];
      x.d1[i] = t.dnode1[i];
    }
  std::copy(t.S.begin(), t.S.end(), x.S);
  std::copy(t.G.begin(), t.G.end(), x.G);
  // v weights of the nodal points (Gauss-Lobatto rule of the nodes)
  auto &v = op->vpdev;
  v.dofs_x = hd::ipow(nd, mf->dim_x);
  v.dofs_v = hd::ipow(nd, mf->dim_v);
  for (int q = 0; q < v.dofs_v; ++q)
    {
      double w = 1.0;
      int rem = q;
      for (int j = 0; j < mf->dim_v; ++j)
        {
          w *= mf->mesh.h[mf->dim_x + j] * t.node_weights[rem % nd];
          rem /= nd;
        }
      v.wv[q] = w;
    }
  op->density_chunks =
    (int)std::min<long long>(op->n_v_cells, std::max<long long>(1, 4096 / ncx));
  const size_t nxd = (size_t)op->n_x_dofs;
  const size_t nef = (size_t)ncx * hd::ipow(nq, mf->dim_x) * mf->dim_x;
  hipError_t e = hipMalloc(&op->efield, nef * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&op->rho, nxd * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&op->phi, nxd * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&op->work, 3 * nxd * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&op->partial,
                  (size_t)ncx * op->density_chunks * v.dofs_x * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&op->stats, 8 * sizeof(long long));
  if (e == hipSuccess)
    e = hipMemset(op->phi, 0, nxd * sizeof(double));
  if (e == hipSuccess)
    e = hipMemset(op->efield, 0, nef * sizeof(double));
  if (e == hipSuccess)
    e = hipMemset(op->stats, 0, 8 * sizeof(long long));
  if (e == hipSuccess && !mf->collocation &&
      hd::vp_apply_supported(mf->dim_x, mf->dim_v, nd))
    {
      // unit-speed 1D operators per direction and sign (the operator for
      // speed c is c A+ for c >= 0 and c A- for c < 0), then S, S^-1 and
      // the Gauss points
      std::vector<double> tab(hd::vp_apply_table_size(nd), 0.0);
      const int per = 2 * (nd * nd + nd);
      for (int d = 0; d < 4; ++d)
        {
          const auto cp = hd::collapse_direction(t, 1.0, mf->mesh.h[d],
                                                 skew_factor);
          const auto cn = hd::collapse_direction(t, -1.0, mf->mesh.h[d],
                                                 skew_factor);
          double *o = tab.data() + d * per;
          for (int i = 0; i < nd * nd; ++i)
            {
              o[i] = cp.A[i];
              o[nd * nd + nd + i] = -cn.A[i];
            }
          for (int i = 0; i < nd; ++i)
            {
              o[nd * nd + i] = cp.b[i];
              o[2 * nd * nd + nd + i] = -cn.b[i];
            }
          // difference form: A u + b f = A^ (u - u_o 1) + b (f - u_o) with
          // A^ = A + b e_o^T and o the outflow node (N-1 for +, 0 for -);
          // A^ 1 = A 1 + b = 0, so both parts act on differences of nearby
          // values and the mass of a cell changes by exactly b-weighted
          // (f - u_o), which telescopes over a periodic row of cells
          for (int i = 0; i < nd; ++i)
            {
              o[i * nd + nd - 1] += cp.b[i];
              o[nd * nd + nd + i * nd] += -cn.b[i];
            }
        }
      double *o = tab.data() + 4 * per;
      std::copy(t.S.begin(), t.S.begin() + nd * nd, o);
      std::copy(t.Sinv.begin(), t.Sinv.begin() + nd * nd, o + nd * nd);
      std::copy(t.qpoints.begin(), t.qpoints.begin() + nd, o + 2 * nd * nd);
      std::copy(op->vpdev.wv, op->vpdev.wv + nd * nd,
                o + 2 * nd * nd + nd);
      // followed by its FP32 copy (the FP32-mixed kernel's coefficients)
      const size_t n_tab = tab.size();
      std::vector<float> tf(n_tab);
      for (size_t i = 0; i < n_tab; ++i)
        tf[i] = (float)tab[i];
      tab.resize(n_tab + (n_tab + 1) / 2, 0.0);
      std::memcpy(tab.data() + n_tab, tf.data(), n_tab * sizeof(float));
      // then the v-direction factors of the x terms, S^-1 diag(v_d(q)+/-) S
      // along v_d, per v_d cell: [d][cell][sign][i][l] (they depend on the
      // v cell only, so the kernel loads them instead of forming them)
      for (int d = 0; d < 2; ++d)
        for (int c = 0; c < mf->mesh.n[2 + d]; ++c)
          for (int sg = 0; sg < 2; ++sg)
            for (int r = 0; r < nd * nd; ++r)
              {
                const int i = r / nd, l = r % nd;
                const double lo = mf->mesh.lo[2 + d] + c * mf->mesh.h[2 + d];
                double acc = 0.0;
                for (int q = 0; q < nd; ++q)
                  {
                    const double v = lo + t.qpoints[q] * mf->mesh.h[2 + d];
                    const double w = sg == 0 ? std::fmax(v, 0.0)
                                             : std::fmin(v, 0.0);
                    acc += t.Sinv[i * nd + q] * w * t.S[q * nd + l];
                  }
                tab.push_back(acc);
              }
      // per d and v_d cell: bit 0 when the + factor is nonzero (a Gauss
      // point with v_d > 0), bit 1 when the - factor is
      const size_t vfac0 = (size_t)n_tab + (n_tab + 1) / 2;
      const size_t n_vfac = tab.size() - vfac0;
      for (int d = 0; d < 2; ++d)
        for (int c = 0; c < mf->mesh.n[2 + d]; ++c)
          {
            const double lo = mf->mesh.lo[2 + d] + c * mf->mesh.h[2 + d];
            int flag = 0;
            for (int q = 0; q < nd; ++q)
              {
                const double v = lo + t.qpoints[q] * mf->mesh.h[2 + d];
                flag |= (v > 0.0 ? 1 : 0) | (v < 0.0 ? 2 : 0);
              }
            tab.push_back((double)flag);
          }
      // and the FP32 copy of the v factors
      {
        std::vector<float> vf(n_vfac);
        for (size_t i = 0; i < n_vfac; ++i)
          vf[i] = (float)tab[vfac0 + i];
        const size_t at = tab.size();
        tab.resize(at + (n_vfac + 1) / 2, 0.0);
        std::memcpy(tab.data() + at, vf.data(), n_vfac * sizeof(float));
      }
      e = hipMalloc(&op->vtab, tab.size() * sizeof(double));
      if (e == hipSuccess)
        e = hipMemcpy(op->vtab, tab.data(), tab.size() * sizeof(double),
                      hipMemcpyHostToDevice);
      // density partials of a stage's output: one per v cell and x node
      if (e == hipSuccess)
        e = hipMalloc(&op->rho_part, (size_t)op->n_v_cells * op->n_x_dofs *
                                       sizeof(double));
    }
  if (e == hipSuccess && hd::vp_poisson2d_supported(mf->dim_x, nd, ncx))
    {
      std::vector<double> tab(8 * nd * nd);
      hd::vp_poisson2d_table(op->xdev, nd, tab.data());
      e = hipMalloc(&op->ptab, tab.size() * sizeof(double));
      if (e == hipSuccess)
        e = hipMemcpy(op->ptab, tab.data(), tab.size() * sizeof(double),
                      hipMemcpyHostToDevice);
    }
  if (e != hipSuccess)
    {
      hd_advection_destroy(op);
      *out = nullptr;
      return hip_fail(e, "hd_vlasov_poisson_create");
    }
  return ok();
}

int hd_vp_update_field(hd_advection *op, const double *f, void *stream)
{
  if (!op || !f || !op->vp)
    return fail(HD_ERR_INVALID, "hd_vp_update_field: not a Vlasov-Poisson "
                                "operator or null argument");
  return vp_update(op, f, nullptr, as_stream(stream));
}

int hd_vp_set_auto_update(hd_advection *op, int enable)
{
  if (!op || !op->vp)
    return fail(HD_ERR_INVALID, "not a Vlasov-Poisson operator");
  op->vp_auto = enable != 0;
  return ok();
}

int hd_vp_set_poisson_tolerance(hd_advection *op, double rel_tol, int max_it)
{
  if (!op || !op->vp || !(rel_tol > 0) || max_it < 1)
    return fail(HD_ERR_INVALID, "invalid Poisson tolerance");
  op->poisson_tol = rel_tol;
  op->poisson_max_it = max_it;
  return ok();
}

int hd_vp_solver_stats(hd_advection *op, long long *iterations,
                       long long *solves, double *last_rel_residual)
{
  if (!op || !op->vp)
    return fail(HD_ERR_INVALID, "not a Vlasov-Poisson operator");
  long long s[3] = {0, 0, 0};
  const hipError_t e =
    hd_abi::device_sync_copy(s, op->stats, 3 * sizeof(long long),
                             hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return hip_fail(e, "hd_vp_solver_stats");
  if (iterations)
    *iterations = s[0];
  if (solves)
    *solves = s[1];
  if (last_rel_residual)
    std::memcpy(last_rel_residual, &s[2], sizeof(double));
  return ok();
}

int hd_vp_solver_start(hd_advection *op, double *start_rel_residual,
                       double *previous_rel_residual, double *weights,
                       long long *guessed_solves)
{
  if (!op || !op->vp)
    return fail(HD_ERR_INVALID, "not a Vlasov-Poisson operator");
  long long s[8];
  const hipError_t e =
    hd_abi::device_sync_copy(s, op->stats, 8 * sizeof(long long),
                             hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return hip_fail(e, "hd_vp_solver_start");
  if (start_rel_residual)
    std::memcpy(start_rel_residual, &s[3], sizeof(double));
  if (previous_rel_residual)
    std::memcpy(previous_rel_residual, &s[4], sizeof(double));
  if (weights)
    std::memcpy(weights, &s[6], 2 * sizeof(double));
  if (guessed_solves)
    *guessed_solves = s[5];
  return ok();
}

int hd_vp_exact_start(hd_advection *op, int enable)
{
  if (!op || !op->vp)
    return fail(HD_ERR_INVALID, "hd_vp_exact_start: not a Vlasov-Poisson "
                                "operator");
  if (!enable)
    {
      op->fdm_on = false;
      return ok();
    }
  const hd_matrix_free *mf = op->mf;
  const int N = mf->nd;
  if (!op->ptab ||
      !hd::vp_fdm_supported(mf->dim_x, N, mf->mesh.n, op->n_x_cells))
    return fail(HD_ERR_UNSUPPORTED,
                "the exact start needs the 2D tensor-product Poisson solver "
                "(dim_x = 2, k <= 3, at most 1024 x cells) and at most 256 "
                "x nodes per direction");
  if (!op->fdm)
    {
      std::vector<double> tab(8 * N * N);
      hd::vp_poisson2d_table(op->xdev, N, tab.data());
      auto *f = new (std::nothrow) hd::FdmDev();
      if (!f)
        return fail(HD_ERR_NOMEM, "out of host memory");
      const hipError_t e =
        hd::vp_fdm_create(tab.data(), N, mf->mesh.n[0], mf->mesh.n[1], f);
      if (e != hipSuccess)
        {
          delete f;
          return hip_fail(e, "hd_vp_exact_start");
        }
      op->fdm = f;
    }
  if (op->learned)
    {
      const int rc = hd_vp_learned_guess_disable(op);
      if (rc != HD_OK)
        return rc;
    }
  op->fdm_on = true;
  return ok();
}

long long hd_vp_x_n_dofs(const hd_advection *op)
{
  return op && op->vp ? op->n_x_dofs : -1;
}

int hd_vp_potential(hd_advection *op, double *phi, int device_to_host)
{
  if (!op || !op->vp || !phi)
    return fail(HD_ERR_INVALID, "hd_vp_potential: invalid argument");
  const hipError_t e =
    hd_abi::device_sync_copy(device_to_host ? (void *)phi : (void *)op->phi,
              device_to_host ? (void *)op->phi : (void *)phi,
              op->n_x_dofs * sizeof(double),
              device_to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_vp_potential");
}

int hd_vp_field_energy(hd_advection *op, double *out)
{
  // sum_K sum_q (d_e phi)^2 JxW (vlasov_poisson/include/diagnostics.h:
  // 95-141) from the stored field; synchronous, host reduction
  if (!op || !op->vp || !out)
    return fail(HD_ERR_INVALID, "hd_vp_field_energy: invalid argument");
  hd_matrix_free *mf = op->mf;
  const int dx = mf->dim_x, nq = mf->nq, nqx = hd::ipow(nq, dx);
  std::vector<double> ef((size_t)op->n_x_cells * nqx * dx);
  const hipError_t e = hd_abi::device_sync_copy(
    ef.data(), op->efield, ef.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return hip_fail(e, "hd_vp_field_energy");
  for (int k = 0; k < dx; ++k)
    out[k] = 0.0;
  for (long long c = 0; c < op->n_x_cells; ++c)
    for (int q = 0; q < nqx; ++q)
      {
        double jxw = 1.0;
        int rem = q;
        for (int j = 0; j < dx; ++j)
          {
            jxw *= mf->mesh.h[j] * mf->tables.qweights[rem % nq];
            rem /= nq;
          }
        for (int k = 0; k < dx; ++k)
          {
            const double g = ef[((size_t)c * nqx + q) * dx + k];
            out[k] += g * g * jxw;
          }
      }
  return ok();
}

int hd_advection_set_algorithm(hd_advection *op, int algorithm)
{
  if (!op || algorithm < HD_ALGO_AUTO || algorithm > HD_ALGO_FCL)
    return fail(HD_ERR_INVALID, "invalid algorithm");
  if (algorithm == HD_ALGO_FCL && op->vp)
    return fail(HD_ERR_UNSUPPORTED,
                "the face-centric form takes a constant velocity");
  if (algorithm == HD_ALGO_COLLAPSED && !op->collapsed_ok)
    return fail(HD_ERR_UNSUPPORTED,
                "collapsed form needs a periodic mesh in every direction");
  op->algorithm = algorithm;
  return ok();
}

int hd_advection_algorithm(const hd_advection *op, int precision)
{
  if (!op || !valid_precision(precision))
    return fail(HD_ERR_INVALID, "invalid argument");
  const int a = choose_algorithm(op, precision);
  return a < 0 ? fail(HD_ERR_UNSUPPORTED, "no kernel") : a;
}

int hd_advection_apply(hd_advection *op, const void *src, void *dst,
                       int precision, double time, void *stream)
{
  if (!op || !src || !dst)
    return fail(HD_ERR_INVALID, "hd_advection_apply: null argument");
  if (!valid_precision(precision))
    return fail(HD_ERR_INVALID, "invalid precision");
  if (src == dst)
    return fail(HD_ERR_INVALID, "src and dst must not alias");
  hd::StageDev stage{};
  return run_operator(op, 0, precision, src, dst, stage, time,
                      as_stream(stream));
}

namespace
{
// scratch: a vector the last stage may use (the free LSRK register)
int lsrk_stage(hd_advection *op, int precision, const void *ti_in,
               void *ti_out, void *scratch, double *sol, double time,
               double a_dt, double b_dt, void *stream,
               int density = kDensityInput, bool write_partials = false);
} // namespace

int hd_lsrk_stage(hd_advection *op, int precision, const void *ti_in,
                  void *ti_out, double *sol, double time, double a_dt,
                  double b_dt, void *stream)
{
  return lsrk_stage(op, precision, ti_in, ti_out, nullptr, sol, time, a_dt,
                    b_dt, stream);
}

namespace
{
int lsrk_stage(hd_advection *op, int precision, const void *ti_in,
               void *ti_out, void *scratch, double *sol, double time,
               double a_dt, double b_dt, void *stream, int density,
               bool write_partials)
{
  if (!op || !ti_in || !sol)
    return fail(HD_ERR_INVALID, "hd_lsrk_stage: null argument");
  if (!valid_precision(precision))
    return fail(HD_ERR_INVALID, "invalid precision");
  if (ti_out == ti_in || (const void *)sol == ti_in)
    return fail(HD_ERR_INVALID, "ti_out/sol must not alias ti_in");
  hd::StageDev stage{};
  stage.a_dt = a_dt;
  stage.b_dt = b_dt;
  stage.sol = sol;
  stage.ti_out = ti_out;
  stage.scratch = ti_out ? nullptr : scratch;
  stage.rho_part = write_partials && ti_out ? op->rho_part : nullptr;
  return run_operator(op, 1, precision, ti_in, nullptr, stage, time,
                      as_stream(stream), density, sol);
}
} // namespace

int hd_lsrk_coefficients(const char *type, double *a, double *b)
{
  std::vector<double> va, vb;
  if (!rk_table(type, va, vb))
    return fail(HD_ERR_INVALID, std::string("unknown RK scheme ") +
                                  (type ? type : "(null)"));
  if (a)
    std::copy(va.begin(), va.end(), a);
  if (b)
    std::copy(vb.begin(), vb.end(), b);
  return (int)vb.size();
}

int hd_lsrk_step(hd_advection *op, const char *type, int precision,
                 double *sol, void *ti_a, void *ti_b, double time, double dt,
                 void *stream)
{
  if (!op || !sol || !ti_a || !ti_b || ti_a == ti_b)
    return fail(HD_ERR_INVALID, "hd_lsrk_step: invalid registers");
  std::vector<double> a, b;
  if (!rk_table(type, a, b))
    return fail(HD_ERR_INVALID, "unknown RK scheme");
  // the first stage reads a copy of the solution (time_integrators.
  // templates.h:142-157); a Vlasov-Poisson operator on the phase-space
  // kernel takes rho from the same pass over the solution
  const bool fused = vp_kernel_active(op) && op->vp_auto;
  int rc;
  if (fused)
    {
      if (!valid_precision(precision))
        return fail(HD_ERR_INVALID, "invalid precision");
      const hipError_t e = hd::vp_density(
        sol, nullptr, op->vpdev, op->n_x_cells, op->n_v_cells, op->partial,
        op->density_chunks, op->rho, as_stream(stream),
        precision == HD_FP32 ? static_cast<float *>(ti_a) : nullptr,
        precision == HD_FP64 ? static_cast<double *>(ti_a) : nullptr);
      rc = e == hipSuccess ? HD_OK : hip_fail(e, "hd_lsrk_step: density");
    }
  else
    rc = hd_vector_convert(op->mf, sol, ti_a, precision, stream);
  if (rc != HD_OK)
    return rc;
  double sum_b = 0.0;
  void *in = ti_a, *out = ti_b;
  const size_t n = b.size();
  for (size_t s = 0; s < n; ++s)
    {
      const double c = s == 0 ? 0.0 : sum_b + a[s - 1];
      if (s > 0)
        sum_b += b[s - 1];
      const bool last = s == n - 1;
      // Vlasov-Poisson with the phase-space kernel: the first stage takes
      // rho from the FP64 solution, the others from the partials the
      // previous stage wrote of its unrounded output (no pass over f, and
      // in the mixed form no FP32 rounding in the Poisson right-hand side)
      const int density =
        !fused ? kDensityInput : (s == 0 ? kDensityReady : kDensityPartials);
      rc = lsrk_stage(op, precision, in, last ? nullptr : out, out, sol,
                      time + c * dt, last ? 0.0 : a[s] * dt, b[s] * dt, stream,
                      density, fused && !last);
      if (rc != HD_OK)
        return rc;
      std::swap(in, out);
    }
  return ok();
}

int hd_interpolate(hd_matrix_free *mf, int function, const double *params,
                   double time, double *dst, void *stream)
{
  if (!mf || !dst)
    return fail(HD_ERR_INVALID, "hd_interpolate: null argument");
  hd::FuncDev f{};
  f.id = function;
  const int n_params = function == HD_FUNC_LANDAU ? 3 : mf->D;
  for (int d = 0; d < n_params; ++d)
    f.prm[d] = params ? params[d] : 0.0;
  std::copy(mf->tables.nodes.begin(), mf->tables.nodes.end(), f.nodes);
  f.nd = mf->nd;
  f.dim_x = mf->dim_x;
  const long long n = mf->n_dofs;
  const hipError_t e = hd::launch_interpolate(
    dst, mf->mesh, f, time, n, mf->dofs_cell, mf->geo, as_stream(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "hd_interpolate");
}

int hd_add_noise(hd_matrix_free *mf, double *dst, uint64_t seed,
                 double amplitude, void *stream)
{
  if (!mf || !dst)
    return fail(HD_ERR_INVALID, "hd_add_noise: null argument");
  const long long n = mf->n_dofs;
  const hipError_t e =
    hd::launch_noise(dst, seed, amplitude, n, as_stream(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "hd_add_noise");
}

int hd_vector_convert(hd_matrix_free *mf, const double *src, void *dst,
                      int precision, void *stream)
{
  if (!mf || !src || !dst || !valid_precision(precision))
    return fail(HD_ERR_INVALID, "hd_vector_convert: invalid argument");
  const long long n = mf->n_dofs;
  if (precision == HD_FP64)
    {
      const hipError_t e =
        hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice,
                       as_stream(stream));
      return e == hipSuccess ? ok() : hip_fail(e, "hd_vector_convert");
    }
  const hipError_t e = hd::launch_convert_f32(
    src, static_cast<float *>(dst), n, as_stream(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "hd_vector_convert");
}

int hd_norm_and_error(hd_matrix_free *mf, const double *src, int function,
                      const double *params, double time, double *out,
                      void *stream)
{
  if (!mf || !src || !out)
    return fail(HD_ERR_INVALID, "hd_norm_and_error: null argument");
  const long long nc = mf->mesh.n_cells;
  hipError_t e = hipSuccess;
  if (!mf->partial)
    e = hipMalloc(&mf->partial, 2 * nc * sizeof(double));
  if (e != hipSuccess)
    return hip_fail(e, "hd_norm_and_error");
  hd::NormDev nt{};
  std::copy(mf->tables.S.begin(), mf->tables.S.end(), nt.S);
  std::copy(mf->tables.qpoints.begin(), mf->tables.qpoints.end(), nt.qp);
  std::copy(mf->tables.qweights.begin(), mf->tables.qweights.end(), nt.w);
  nt.nd = mf->nd;
  nt.nq = mf->nq;
  hd::FuncDev f{};
  f.id = function;
  for (int d = 0; d < mf->D; ++d)
    f.prm[d] = params ? params[d] : 0.0;
  f.nd = mf->nd;
  f.dim_x = mf->dim_x;
  const int vol = hd::ipow(std::max(mf->nd, mf->nq), mf->D);
  const size_t lds = (2 * (size_t)vol + 512) * sizeof(double);
  if (lds > 160 * 1024)
    return fail(HD_ERR_UNSUPPORTED, "cell too large for hd_norm_and_error");
  e = hd::launch_norm(src, mf->mesh, nt, f, time, vol, mf->partial, mf->geo,
                      as_stream(stream));
  if (e != hipSuccess)
    return hip_fail(e, "hd_norm_and_error");
  std::vector<double> host(2 * nc);
  e = hipMemcpyAsync(host.data(), mf->partial, 2 * nc * sizeof(double),
                     hipMemcpyDeviceToHost, as_stream(stream));
  if (e == hipSuccess)
    e = hipStreamSynchronize(as_stream(stream));
  if (e != hipSuccess)
    return hip_fail(e, "hd_norm_and_error");
  double s0 = 0.0, s1 = 0.0;
  for (long long c = 0; c < nc; ++c)
    {
      s0 += host[2 * c];
      s1 += host[2 * c + 1];
    }
  out[0] = std::sqrt(s0);
  out[1] = std::sqrt(s1);
  return ok();
}

int hd_vp_moments(hd_matrix_free *mf, const double *f, double *out,
                  void *stream)
{
  // mass, ||f||, kinetic energy, momentum (3) at the quadrature points
  // (vlasov_poisson/include/diagnostics.h:28-87); synchronous
  if (!mf || !f || !out)
    return fail(HD_ERR_INVALID, "hd_vp_moments: null argument");
  const long long nc = mf->mesh.n_cells;
  double *partial = nullptr;
  hipError_t e = hipMalloc(&partial, 6 * nc * sizeof(double));
  if (e != hipSuccess)
    return hip_fail(e, "hd_vp_moments");
  hd::NormDev nt{};
  std::copy(mf->tables.S.begin(), mf->tables.S.end(), nt.S);
  std::copy(mf->tables.qpoints.begin(), mf->tables.qpoints.end(), nt.qp);
  std::copy(mf->tables.qweights.begin(), mf->tables.qweights.end(), nt.w);
  nt.nd = mf->nd;
  nt.nq = mf->nq;
  const int vol = hd::ipow(std::max(mf->nd, mf->nq), mf->D);
  e = hd::vp_moments(f, mf->mesh, nt, mf->dim_x, vol, partial,
                     as_stream(stream));
  std::vector<double> host(6 * nc);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host.data(), partial, 6 * nc * sizeof(double),
                       hipMemcpyDeviceToHost, as_stream(stream));
  if (e == hipSuccess)
    e = hipStreamSynchronize(as_stream(stream));
  (void)hipFree(partial);
  if (e != hipSuccess)
    return hip_fail(e, "hd_vp_moments");
  for (int k = 0; k < 6; ++k)
    out[k] = 0.0;
  for (long long cc = 0; cc < nc; ++cc)
    for (int k = 0; k < 6; ++k)
      out[k] += host[6 * cc + k];
  out[1] = std::sqrt(out[1]);
  return ok();
}

double hd_critical_time_step(const hd_matrix_free *mf, const double *velocity)
{
  if (!mf || !velocity)
    return -1.0;
  double vx = 0.0, vv = 0.0;
  if (mf->deformed)
    {
      // cfl.h:56-90: the largest |(J^{-T} a)_j| at the Gauss points
      // (n_points) of each space's cells
      hd::Tables1D g;
      if (!g.build(mf->degree, mf->nq, false))
        return -1.0;
      vx = hd::space_max_speed(mf->xmap, mf->nq, g.qpoints.data(), velocity);
      vv = hd::space_max_speed(mf->vmap, mf->nq, g.qpoints.data(),
                               velocity + mf->dim_x);
    }
  else
    for (int d = 0; d < mf->D; ++d)
      {
        const double v = std::fabs(velocity[d] / mf->mesh.h[d]);
        if (d < mf->dim_x)
          vx = std::max(vx, v);
        else
          vv = std::max(vv, v);
      }
  return std::min(1.0 / vx, 1.0 / vv);
}

int hd_device_malloc(void **ptr, uint64_t bytes)
{
  if (!ptr)
    return fail(HD_ERR_INVALID, "null pointer");
  const hipError_t e = hipMalloc(ptr, bytes);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_device_malloc");
}

int hd_device_free(void *ptr)
{
  const hipError_t e = hipFree(ptr);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_device_free");
}

int hd_memcpy_htod(void *dst, const void *src, uint64_t bytes)
{
  const hipError_t e =
    hd_abi::device_sync_copy(dst, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_memcpy_htod");
}

int hd_memcpy_dtoh(void *dst, const void *src, uint64_t bytes)
{
  const hipError_t e =
    hd_abi::device_sync_copy(dst, src, bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_memcpy_dtoh");
}

int hd_device_synchronize(void)
{
  const hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? ok() : hip_fail(e, "hd_device_synchronize");
}

} // extern "C"
