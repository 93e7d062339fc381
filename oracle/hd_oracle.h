/*
 * hd_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of
 * hyper.deal's matrix-free phase-space DG advection operator (cell integrals,
 * upwind face integrals, inverse mass matrix), of its low-storage Runge-Kutta
 * integrators and of the interpolation / L2-norm helpers.  It exists to check
 * the HIP product path and to serve as the timed CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Data layout (shared with the product so results compare element-wise):
 *   - phase-space cells are numbered lexicographically over the D = dim_x +
 *     dim_v Cartesian directions, x0 fastest (x-cells before v-cells, as in
 *     hyper.deal's lid = lid_x + lid_v * n_x, matrix_free.templates.h:552-562);
 *   - each cell stores its (k+1)^D nodal values contiguously, nodes at the
 *     Gauss-Lobatto points, in-cell index lexicographic with direction 0
 *     fastest (i_x + (k+1)^dim_x * i_v, read_write_operation.h:143-174).
 */
#ifndef HD_ORACLE_H
#define HD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hdo_problem hdo_problem;

/* Analytic functions the oracle can interpolate / compare against.
 * params: the advection vector a (D entries) where the function moves with it. */
enum hdo_function
{
  HDO_FUNC_ZERO = 0,
  HDO_FUNC_HYPERRECTANGLE = 1, /* sin(2 pi (x0-a0 t)) prod_{d>0} cos(2 pi (x_d - a_d t)) */
  HDO_FUNC_CONSTANT_ONE = 2,
  HDO_FUNC_SIN_COS = 3,        /* sin(pi x0) prod_{d>0} cos(pi x_d)                      */
  HDO_FUNC_PRODUCT_SINES = 4,  /* prod_d sin(2 pi (x_d - a_d t))                          */
  HDO_FUNC_LANDAU = 5          /* (1 + eps prod_{d<np} cos(kx x_d)) exp(-|v|^2/2)/(2pi)^{dv/2}, params eps, kx, np */
};

/* Boundary treatment of non-periodic directions (inhomogeneous Dirichlet uses
 * the analytic function bc_func with the same params). */
enum hdo_boundary
{
  HDO_BC_HOMOGENEOUS = 0,
  HDO_BC_INHOMOGENEOUS = 1
};

/*
 * Create a Cartesian phase-space problem.
 *   n_cells[D], lo[D], hi[D]: cells per direction and the box.
 *   n_points: 1D quadrature points (Gauss, or Gauss-Lobatto when collocation).
 *   a[D]: constant transport velocity.
 *   func_id/func_params: analytic function for Dirichlet data and errors;
 *   func_params holds D values (3 for HDO_FUNC_LANDAU), or NULL for a.
 */
hdo_problem *hdo_create(int dim_x, int dim_v, int degree, int n_points,
                        int collocation, const int *n_cells, const double *lo,
                        const double *hi, int periodic_x, int periodic_v,
                        double skew, const double *a, int bc_kind, int func_id,
                        const double *func_params);
void hdo_destroy(hdo_problem *p);
long long hdo_n_dofs(const hdo_problem *p);
long long hdo_n_cells(const hdo_problem *p);

/* dst = M^{-1} (cell + face integrals)(src) at time t; use_fcl selects the
 * face-centric loop instead of the element-centric one. */
int hdo_apply(hdo_problem *p, const double *src, double *dst, double t,
              int use_fcl);

/* ECL apply on the listed cells only (dst written for those cells). */
int hdo_apply_cells(hdo_problem *p, const double *src, double *dst, double t,
                    const long long *cell_ids, int n);

/* Interpolate the problem's function (func_id) at time t into dst. */
int hdo_interpolate(hdo_problem *p, double t, double *dst);

/* out[0] = ||u||_L2, out[1] = ||u - f(t)||_L2 with the problem's function. */
int hdo_norm_and_error(hdo_problem *p, const double *src, double t,
                       double *out);

/* dt_crit = min over x and v of 1/max |J^{-T} a|. */
double hdo_critical_time_step(const hdo_problem *p);
/* Deformed mesh (grid_generator.cc:110-230, DeformedCubeManifold, applied
 * as the exact map): space x maps x^ -> x^ + dx prod_j sin(fx pi (x^_j -
 * lo_j) / (hi_j - lo_j)) (1, ..., 1), space v likewise with (dv, fv);
 * 0, 0 restores the Cartesian mesh.  Not with Vlasov-Poisson. */
int hdo_set_deformation(hdo_problem *p, double deformation_x, int frequency_x,
                        double deformation_v, int frequency_v);
/* The mesh map at reference point xi of the cell with multi-index
 * cell_index: physical point x (D), inverse Jacobian d xi_i / d x_j
 * (D x D, row-major) and det of d x / d xi; any output may be NULL. */
int hdo_geometry(const hdo_problem *p, const int *cell_index,
                 const double *xi, double *x, double *inverse_jacobian,
                 double *det);

/* One low-storage Runge-Kutta step (rk33, rk45, rk47, rk59) of du/dt = op(u),
 * op = apply (ECL when use_ecl else FCL).  ki, ti: caller-owned registers. */
int hdo_lsrk_step(hdo_problem *p, const char *rk_type, double *sol, double *ki,
                  double *ti, double t, double dt, int use_ecl);

/* Scalar ODE y' = y sin^2(t) integrated with LSRK (time_integrators_02). */
double hdo_lsrk_ode(const char *rk_type, double dt, int n_steps, double y0);
int hdo_lsrk_n_stages(const char *rk_type);
/* Copies the scheme's a (n_stages-1) and b (n_stages) coefficients. */
int hdo_lsrk_coefficients(const char *rk_type, double *a, double *b);

/*
 * The advection example end to end (examples/advection): interpolate f0 at
 * the Gauss-Lobatto nodes, dt from the CFL rule, time loop with LSRK, and the
 * (t, ||u||, ||u-u_ex||) diagnostics rows printed at every diagnostics tick.
 * Returns the number of rows written (3 doubles each), or -1 on error.
 */
int hdo_run_advection(hdo_problem *p, const char *rk_type, double time_step,
                      double start_time, double final_time, double cfl,
                      double diag_tick, int use_ecl, int max_steps,
                      double *rows, int max_rows, double *dt_used);

/* 1D tables: nodes (k+1 Gauss-Lobatto), quadrature points/weights, S (nq x
 * nd), D (nq x nq), face values F (2 x nq), Sinv (nd x nq), on [0,1]. */
int hdo_tables(int degree, int n_points, int collocation, double *nodes,
               double *qpoints, double *qweights, double *S, double *D,
               double *F, double *Sinv);

/* Seeded inputs: counter-based SplitMix64 in [-1, 1): value for counter i. */
double hdo_splitmix_uniform(unsigned long long seed, unsigned long long i);
/* dst += amplitude * U(-1,1) noise keyed by the global DoF index. */
int hdo_add_noise(const hdo_problem *p, double *dst, unsigned long long seed,
                  double amplitude);

/* ---- Vlasov-Poisson (examples/vlasov_poisson), requires dim_x == dim_v ----
 * The x-space vectors (rho, b, phi) hold (k+1)^dim_x nodal values per x cell,
 * x cells lexicographic. */
int hdo_set_vlasov_poisson(hdo_problem *p, int enable);
long long hdo_x_n_dofs(const hdo_problem *p);
/* rho = integral of f over v (Gauss-Lobatto v rule on the nodal values) */
int hdo_vp_density(hdo_problem *p, const double *f, double *rho);
/* rho -= mean; b = -M rho; b -= mean */
int hdo_vp_rhs(hdo_problem *p, double *rho, double *b);
/* dst = A (src - mean(src)), A the periodic SIPG Laplacian */
int hdo_poisson_apply(hdo_problem *p, const double *src, double *dst);
/* CG on A phi = b from the given phi; returns iterations */
int hdo_poisson_solve(hdo_problem *p, const double *b, double *phi,
                      double rel_tol, int max_it);
/* store grad phi at the x quadrature points as the v-velocity */
int hdo_vp_set_potential(hdo_problem *p, const double *phi);
int hdo_vp_efield(const hdo_problem *p, double *out);
/* sum (d_e phi)^2 JxW per x component, from the stored field */
int hdo_vp_energy(hdo_problem *p, double *out);
/* mass, l2 norm, kinetic energy, momentum[3] of f */
int hdo_vp_diagnostics(hdo_problem *p, const double *f, double *out);
/* the Vlasov-Poisson example end to end; rows of 1 + dim_x + 6 values */
int hdo_run_vlasov_poisson(hdo_problem *p, const char *rk_type,
                           double time_step, double start_time,
                           double final_time, double cfl, double diag_tick,
                           double poisson_tol, int max_steps, double *rows,
                           int max_rows, double *dt_used, int *cg_iterations);

/* Number of OpenMP threads the oracle uses. */
int hdo_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
