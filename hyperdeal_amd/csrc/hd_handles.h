// hd_handles.h -- the opaque handles of the C ABI and the error helpers
// shared by the translation units that implement it (host code only).
#pragma once

#include "../../include/hyperdeal_amd.h"
#include "hd_cnn.h"
#include "hd_device.h"
#include "hd_geometry.h"
#include "hd_tables.h"

#include <string>

namespace hd_abi
{
// message of the last failure on this thread (hd_last_error)
extern thread_local std::string g_last_error;

inline int fail(int code, const std::string &msg)
{
  g_last_error = msg;
  return code;
}

inline int hip_fail(hipError_t e, const char *where)
{
  return fail(HD_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

inline int ok()
{
  return HD_OK;
}

inline hipStream_t as_stream(void *s)
{
  return reinterpret_cast<hipStream_t>(s);
}

// A copy of operator state between host and device for the entry points
// that take no stream (statistics, the potential, the field energy, network
// parameters).  The state is written by kernels on the caller's streams,
// which may be non-blocking (they do not wait for the null stream, nor it for
// them), so the whole device is synchronised before the copy; the copy itself
// is blocking, so work issued afterwards on any stream sees its result.
inline hipError_t device_sync_copy(void *dst, const void *src, size_t bytes,
                                   hipMemcpyKind kind)
{
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(dst, src, bytes, kind);
  if (e == hipSuccess && kind != hipMemcpyDeviceToHost)
    e = hipDeviceSynchronize();
  return e;
}
} // namespace hd_abi

struct hd_matrix_free
{
  int dim_x = 0, dim_v = 0, D = 0, degree = 0, nd = 0, nq = 0;
  int collocation = 0;
  bool periodic[6] = {true, true, true, true, true, true};
  hd::MeshDev mesh{};
  hd::Tables1D tables;
  long long n_dofs = 0;
  int dofs_cell = 0;
  double *partial = nullptr; // norm partial sums (2 per cell)
  // deformed mesh: the maps of the two spaces, |det J| w per space, cell and
  // quadrature point (device, x then v) and their view for the kernels
  bool deformed = false;
  hd::SpaceMap xmap, vmap;
  double *geo_jw = nullptr;
  hd::DeformDev geo{};
  int n_ops = 0; // operators created on this object and not destroyed
};

// a standalone network (hd_cnn_*), also the one inside a learned guess
struct hd_cnn
{
  hd::CnnNet net;
};

// the learned initial guess of a Vlasov-Poisson operator's potential solve:
// the network, a ring of (right-hand side, solution) training images taken
// from the operator's own converged solves -- the residual of the previous
// potential and the correction CG added to it -- and the training schedule
struct hd_learned_guess
{
  hd_cnn cnn;
  int n0 = 0, n1 = 0, N = 0; // x cells per direction, nodes per direction
  long long plane = 0;       // pixels per image
  int capacity = 0, count = 0, head = 0;
  float *ring_in = nullptr, *ring_tgt = nullptr;
  float *img_out = nullptr, *batch_in = nullptr, *batch_tgt = nullptr;
  double *scale = nullptr, *scale_t = nullptr, *guess = nullptr;
  double *phi_prev = nullptr; // x-space DoFs
  uint64_t seed = 0, draws = 0;
  long long solves = 0, guessed = 0;
  long long dense_solves = 0; // the first solves all train (then train_every)
  int train_every = 1, train_batch = 8, train_steps = 1, min_pairs = 4;
  float lr = 1e-3f;
  // training runs on its own stream while the solve's CG (one workgroup)
  // runs on the caller's: forked after the guess, joined before the next
  hipStream_t train_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_done = nullptr;
  bool pending = false; // training enqueued and not yet joined
};

struct hd_advection
{
  hd_matrix_free *mf = nullptr;
  double a[6] = {0, 0, 0, 0, 0, 0};
  double skew = 0.0;
  int bc_kind = HD_BC_INHOMOGENEOUS;
  int bc_function = HD_FUNC_ZERO;
  double bc_params[8] = {0};
  int algorithm = HD_ALGO_AUTO;
  bool collapsed_ok = false;
  bool dirichlet = false;      // some direction has Dirichlet boundaries
  bool split = false;          // the collapsed form runs split (big cells)
  // a free vector (n_dofs doubles) for the last stage of the split form
  // (its first part) and of the FCL form (the assembled integrals)
  void *split_scratch = nullptr;
  double *coef64 = nullptr;
  float *coef32 = nullptr;
  hd::InflowDev inflow{};
  hd::QuadDev quad{};
  hd::MassDev mass{};          // inverse mass matrix of the FCL form
  // deformed mesh: the operator's per-space tables (J^{-1} a and JxW at the
  // cell points, (a . n) JxW_face at the face points) and their view
  double *geo_tab = nullptr;
  hd::DeformDev geo{};
  bool counted = false; // counted in mf->n_ops
  // Vlasov-Poisson state (velocity (v, grad phi) from the Poisson solve)
  bool vp = false;
  bool vp_auto = true;        // stages recompute the field from their input
  double poisson_tol = 1e-10; // CG relative residual
  int poisson_max_it = 100000;
  hd::XDev xdev{};
  hd::VpDev vpdev{};
  long long n_x_cells = 0, n_v_cells = 0, n_x_dofs = 0;
  int density_chunks = 1;
  double *efield = nullptr, *rho = nullptr, *phi = nullptr, *work = nullptr,
         *partial = nullptr;
  long long *stats = nullptr; // solver statistics, see hd::vp_poisson
  double *ptab = nullptr;     // 2D SIPG table of the fast solver, or null
  double *vtab = nullptr;     // tables of the 2+2 phase-space apply, or null
  double *rho_part = nullptr; // density partials written by a stage
  hd_learned_guess *learned = nullptr;
  // the exact start of the 2D potential solve (fast diagonalisation), or null
  hd::FdmDev *fdm = nullptr;
  bool fdm_on = false;
};

namespace hd_abi
{
// the learned guess around one potential solve: before it (the residual of
// the previous potential as an image, network inference, the guess in
// x-space DoFs; returns the guess or null) and after it (training pair,
// scheduled training steps); rho must be current
hipError_t learned_before_solve(hd_advection &op, const double **guess,
                                hipStream_t s);
hipError_t learned_after_solve(hd_advection &op, hipStream_t s);
void learned_release(hd_learned_guess *lg);
} // namespace hd_abi
