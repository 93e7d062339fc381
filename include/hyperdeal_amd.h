/*
 * hyperdeal_amd.h -- C ABI of the MI355X (gfx950) phase-space DG advection
 * operator: the drop-in boundary behind hyperdeal::MatrixFree's loops and
 * hyperdeal::advection::AdvectionOperation.
 *
 * Conventions
 *   - Status: every entry returns HD_OK (0) or a negative HD_ERR_* code; the
 *     message of the last failure on the calling thread is hd_last_error().
 *     No C++ exception crosses this boundary.  (The reference raises
 *     dealii::ExceptionBase through AssertThrow, e.g.
 *     base/time_integrators.templates.h:87-88.)
 *   - Vectors are caller-owned device pointers (hipMalloc'd, or any HIP
 *     device allocation) holding n_dofs values, cell-contiguous: cells are
 *     numbered lexicographically over the D = dim_x + dim_v directions with
 *     x0 fastest, each cell holds its (k+1)^D nodal values (Gauss-Lobatto
 *     nodes) lexicographically with direction 0 fastest -- the in-cell order
 *     of hyper.deal (i_x + (k+1)^dim_x i_v, matrix_free/
 *     read_write_operation.h:143-174) and its cell order lid_x + lid_v n_x
 *     (matrix_free/matrix_free.templates.h:552-562) for a subdivided box.
 *   - stream: a hipStream_t passed as void*; NULL is the default stream.  All
 *     work on one handle is stream ordered; a handle is not thread safe
 *     (hyper.deal's evaluators are not re-entrant either,
 *     matrix_free/fe_evaluation_base.h:93).
 */
#ifndef HYPERDEAL_AMD_H
#define HYPERDEAL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HD_OK 0
#define HD_ERR_INVALID -1     /* bad argument */
#define HD_ERR_UNSUPPORTED -2 /* valid but not implemented for this shape */
#define HD_ERR_HIP -3         /* HIP runtime failure */
#define HD_ERR_NOMEM -4

typedef enum hd_precision
{
  HD_FP64 = 0, /* vectors and arithmetic in double */
  HD_FP32 = 1  /* operator vectors and arithmetic in float */
} hd_precision;

typedef enum hd_algorithm
{
  HD_ALGO_AUTO = 0,       /* collapsed where exact, else quadrature */
  HD_ALGO_QUADRATURE = 1, /* the reference's sum-factorised quadrature steps */
  HD_ALGO_COLLAPSED = 2,  /* exact tensor factorisation (constant velocity,
                             periodic Cartesian mesh) */
  HD_ALGO_FCL = 3         /* the reference's face-centric loop: cell, face
                             and boundary integrals assembled once per cell
                             and face, then the inverse mass matrix
                             (MatrixFree::loop + cell_loop, matrix_free.
                             templates.h:1417-1465, 1634-1746; advection_
                             operation.h:159-183, 573-1131) */
} hd_algorithm;

/* Boundary data of non-periodic directions (operators/advection/
 * boundary_descriptor.h:31-115): u+ = -u- (homogeneous) or -u- + 2 g. */
typedef enum hd_boundary
{
  HD_BC_HOMOGENEOUS = 0,
  HD_BC_INHOMOGENEOUS = 1
} hd_boundary;

/* Analytic functions for initial data, Dirichlet data and errors. */
typedef enum hd_function
{
  HD_FUNC_ZERO = 0,
  HD_FUNC_HYPERRECTANGLE = 1, /* sin(2pi(x0-a0 t)) prod cos(2pi(xd-ad t)),
                                 examples/advection/cases/hyperrectangle.h:33-55 */
  HD_FUNC_CONSTANT_ONE = 2,
  HD_FUNC_SIN_COS = 3,        /* sin(pi x0) prod cos(pi xd), tests_functions.h */
  HD_FUNC_PRODUCT_SINES = 4,  /* prod sin(2pi(xd - ad t)) */
  HD_FUNC_LANDAU = 5          /* (1 + eps prod_{d<np} cos(k x_d)) exp(-|v|^2/2)
                                 / (2pi)^{dim_v/2}, params (eps, k, np),
                                 vlasov_poisson/cases/hyperrectangle.h:44-58 */
} hd_function;

/* A Cartesian phase-space box: dim_x + dim_v directions (x first). */
typedef struct hd_mesh_desc
{
  int dim_x, dim_v;   /* 1..3 each */
  int n_cells[6];     /* cells per direction */
  double lo[6], hi[6];
  int periodic_x, periodic_v;
} hd_mesh_desc;

typedef struct hd_matrix_free hd_matrix_free;
typedef struct hd_advection hd_advection;

/* Thread-local message of the last failing call ("" if none). */
const char *hd_last_error(void);

/* Library version (major, minor). */
int hd_version(int *major, int *minor);

/*
 * Phase-space matrix-free data: replaces hyperdeal::MatrixFree::reinit
 * (matrix_free/matrix_free.templates.h:860-1235) together with the two
 * dealii::MatrixFree objects and ShapeInfo::reinit (matrix_free/
 * shape_info.h:109-231).  n_points is the 1D quadrature size (Gauss, or
 * Gauss-Lobatto when collocation != 0, then n_points == degree + 1).
 */
int hd_matrix_free_create(const hd_mesh_desc *mesh, int degree, int n_points,
                          int collocation, hd_matrix_free **out);
/* Fails (HD_ERR_INVALID, nothing freed) while operators created on mf are
 * alive: destroy them first. */
int hd_matrix_free_destroy(hd_matrix_free *mf);
/*
 * Deform the mesh: replaces subdivided_hyper_rectangle(...,
 * with_internal_deformation = true) with DeformedCubeManifold
 * (source/grid/grid_generator.cc:110-230, 240-290).  Each space is mapped by
 *   x^ -> x^ + deformation prod_j sin(frequency pi (x^_j - lo_j) /
 *                                     (hi_j - lo_j)) (1, ..., 1)
 * (the reference's defaults: deformation 0.1, frequency 2), used as the
 * exact map of every cell (the reference interpolates it with its
 * MappingQ).  Operators then run the quadrature (ECL) or FCL kernels with
 * per-point Jacobians, JxW and normals; interpolation, norms and the
 * critical time step use the mapped points.  Call before creating
 * operators on mf (HD_ERR_INVALID otherwise); 0, 0 restores the Cartesian
 * mesh.  Not for Vlasov-Poisson; the FCL form needs n_points = degree + 1.
 */
int hd_matrix_free_set_deformation(hd_matrix_free *mf, double deformation_x,
                                   int frequency_x, double deformation_v,
                                   int frequency_v);
/* n_dofs, n_cells, dofs per cell: MatrixFree::initialize_dof_vector sizes
 * (matrix_free.templates.h:1367-1413). */
long long hd_matrix_free_n_dofs(const hd_matrix_free *mf);
long long hd_matrix_free_n_cells(const hd_matrix_free *mf);
int hd_matrix_free_dofs_per_cell(const hd_matrix_free *mf);

/*
 * Advection operator: replaces AdvectionOperation::reinit
 * (operators/advection/advection_operation.h:97-129) with a constant
 * transport velocity (ConstantVelocityFieldView, velocity_field_view.h:
 * 64-166), skew factor (advection_operation_parameters.h:29-42) and the
 * boundary descriptor.  bc_params: the function's D parameters (its
 * velocity), may be NULL.
 */
int hd_advection_create(hd_matrix_free *mf, const double *velocity,
                        double skew_factor, int bc_kind, int bc_function,
                        const double *bc_params, hd_advection **out);
int hd_advection_destroy(hd_advection *op);
/* Force an algorithm (HD_ALGO_*); HD_ERR_UNSUPPORTED if not applicable.
 * HD_ALGO_COLLAPSED: Cartesian boxes, any n_points >= degree + 1;
 * Dirichlet directions through a boundary correction after the kernel
 * (inhomogeneous data: at most 7776 face points, n_points^(dim-1)).
 * HD_ALGO_QUADRATURE: any boundary, n_points = degree + 1 or degree + 2.
 * HD_ALGO_FCL: as HD_ALGO_QUADRATURE, constant velocity; a stage without
 * ti_out assembles in a vector the operator allocates on first use.
 * (AdvectionOperation chooses ECL or FCL with Matrixfree/UseECL,
 * advection_operation.h:157-204.) */
int hd_advection_set_algorithm(hd_advection *op, int algorithm);
/* The algorithm apply() runs with the current settings and precision. */
int hd_advection_algorithm(const hd_advection *op, int precision);

/*
 * dst = M^{-1} [ cell integrals + upwind face integrals ](src) at time t:
 * replaces AdvectionOperation::apply (advection_operation.h:134-209), i.e.
 * MatrixFree::loop_cell_centric with local_apply_advect_and_inverse_mass_
 * matrix (matrix_free.templates.h:1494-1581, advection_operation.h:221-566).
 * Periodic neighbours are read straight from src.  src and dst hold double
 * (HD_FP64) or float (HD_FP32) values and must not overlap.
 */
int hd_advection_apply(hd_advection *op, const void *src, void *dst,
                       int precision, double time, void *stream);

/*
 * One fused low-storage Runge-Kutta stage: K = apply(ti_in, time), then
 *   sol += b_dt K;  ti_out = sol_old + a_dt K  (skipped when ti_out == NULL)
 * -- the stage of LowStorageRungeKuttaIntegrator::perform_time_step
 * (base/time_integrators.templates.h:103-138) fused into the operator sweep.
 * sol is FP64; ti_in/ti_out are `precision`.  ti_out must differ from ti_in.
 */
int hd_lsrk_stage(hd_advection *op, int precision, const void *ti_in,
                  void *ti_out, double *sol, double time, double a_dt,
                  double b_dt, void *stream);

/* Coefficients of rk33/rk45/rk47/rk59 (time_integrators.templates.h:32-84):
 * a[n_stages-1], b[n_stages].  Returns the number of stages or an error. */
int hd_lsrk_coefficients(const char *type, double *a, double *b);

/*
 * A whole LSRK time step (time_integrators.templates.h:91-184) with the
 * stages fused into the apply: ti_a and ti_b are two `precision` registers
 * (ping-pong, so neighbours are never read while written).
 */
int hd_lsrk_step(hd_advection *op, const char *type, int precision,
                 double *sol, void *ti_a, void *ti_b, double time, double dt,
                 void *stream);

/*
 * Batches of independent advection operators.  The reference has no
 * counterpart (it runs one operator per MPI rank); a batch replaces many
 * calls of AdvectionOperation::apply (advection_operation.h:134-209) and
 * LowStorageRungeKuttaIntegrator::perform_time_step (time_integrators.
 * templates.h:91-184) on small meshes by one launch per degree and apply /
 * stage (the operators of one degree share a kernel with the block shape of
 * that degree's single-operator kernel).  The
 * operators must use the collapsed kernel on 2+2 periodic meshes (degree
 * 1..5, meshes and degrees may differ); at most HD_BATCH_MAX of them.  Array
 * arguments hold one pointer per operator, in the order of hd_batch_create.
 * Graphs must be destroyed before their batch.
 */
#define HD_BATCH_MAX 64
typedef struct hd_batch hd_batch;
typedef struct hd_graph hd_graph;

int hd_batch_create(hd_advection *const *ops, int n_ops, hd_batch **out);
int hd_batch_destroy(hd_batch *batch);
int hd_batch_size(const hd_batch *batch);
/* dst[i] = apply_i(src[i]) for every operator, one launch */
int hd_batch_apply(hd_batch *batch, const void *const *src, void *const *dst,
                   int precision, void *stream);
/* one LSRK step of every operator: sol[i] FP64, ti_a[i] != ti_b[i] stage
 * registers of `precision`; one copy launch and one launch per degree and
 * stage */
int hd_batch_lsrk_step(hd_batch *batch, const char *type, int precision,
                       double *const *sol, void *const *ti_a,
                       void *const *ti_b, double dt, void *stream);
/* Work split: every operator's cells are cut into eight contiguous parts,
 * one per XCD, and each XCD's blocks walk its parts interleaved (a task: one
 * block's share of one part).  The autotuner times each degree's batched
 * apply on src -> dst with tasks capped at 1, 4 or 16 cell groups or
 * uncapped, and keeps the fastest cap per degree.  Synchronous on
 * `stream`. */
int hd_batch_autotune(hd_batch *batch, int precision, const void *const *src,
                      void *const *dst, void *stream);
/* the mean cells per task of a degree in the current work split (a task: one
 * block's share of one fragment), or an error */
int hd_batch_cells_per_task(const hd_batch *batch, int degree, int precision);
/* caps a degree's tasks at `cells` cells (rounded up to whole cell groups) */
int hd_batch_set_cells_per_task(hd_batch *batch, int degree, int precision,
                                int cells);
/* a hipGraph of hd_batch_lsrk_step with these vectors and dt: each launch
 * advances every operator by dt */
int hd_batch_graph_create(hd_batch *batch, const char *type, int precision,
                          double *const *sol, void *const *ti_a,
                          void *const *ti_b, double dt, hd_graph **out);
int hd_graph_launch(hd_graph *graph, void *stream);
int hd_graph_destroy(hd_graph *graph);

/*
 * Vlasov-Poisson operator (examples/vlasov_poisson/include/application.h:
 * 516-602): the advection apply with the phase-space velocity (v, grad phi)
 * (vlasov_poisson/include/velocity_field_view.h:28-209), where phi solves
 * the periodic SIPG Poisson problem -lap phi = -(rho - mean rho)
 * (poisson.h:52-312) for rho = int f dv (numerics/vector_tools.h:230-315).
 * Needs dim_x == dim_v, a periodic box, n_points = degree + 1.  The apply
 * runs in FP64 or in the FP32-mixed form (FP32 vectors and contractions;
 * the density, the potential solve and the LSRK update of the solution stay
 * FP64).  With auto update on (the default) every apply / stage first
 * recomputes rho, phi (FP64 conjugate gradients from the previous phi to
 * poisson_tol relative residual) and grad phi from its input, as the
 * reference's RK right-hand side does.
 */
int hd_vlasov_poisson_create(hd_matrix_free *mf, double skew_factor,
                             double poisson_tol, hd_advection **out);
/* rho, phi and grad phi from f (FP64 phase-space vector) */
int hd_vp_update_field(hd_advection *op, const double *f, void *stream);
int hd_vp_set_auto_update(hd_advection *op, int enable);
int hd_vp_set_poisson_tolerance(hd_advection *op, double rel_tol, int max_it);
/* accumulated CG iterations and solves, last relative residual.  This and
 * the other entry points below that take no stream (hd_vp_solver_start,
 * hd_vp_potential, hd_vp_field_energy, hd_cnn_params, hd_cnn_last_loss,
 * hd_vp_learned_guess_stats, hd_memcpy_htod / dtoh) synchronise the whole
 * device before they read or write operator state, so they see the work of
 * every stream, non-blocking ones included, and work issued after them sees
 * what they wrote. */
int hd_vp_solver_stats(hd_advection *op, long long *iterations,
                       long long *solves, double *last_rel_residual);
/* the start of the last solve (device-synchronising): its relative residual, that of the
 * previous phi alone, the weights (previous phi, guess) of the start, and
 * the number of solves that started from a learned guess */
int hd_vp_solver_start(hd_advection *op, double *start_rel_residual,
                       double *previous_rel_residual, double *weights,
                       long long *guessed_solves);
/* Exact start of the potential solve (enable != 0; off by default, when CG
 * starts from the previous phi as in the reference, vlasov_poisson/include/
 * application.h:575-576).  No counterpart in the reference.  On a periodic
 * 2D x mesh the SIPG system separates into 1D operators per direction; their
 * generalised eigenvectors (computed once, on the host, at the first enable)
 * solve it with four dense products of (cells x nodes per direction)-sized
 * matrices, and CG starts from that solution and checks it to the same
 * tolerance (typically zero or one iteration).  dim_x == 2, the 2D
 * tensor-product solver (k <= 3, at most 1024 x cells), at most 256 x nodes
 * per direction.  Enabling it disables a learned guess. */
int hd_vp_exact_start(hd_advection *op, int enable);
/* x-space DoFs of phi / rho: (k+1)^dim_x per x cell */
long long hd_vp_x_n_dofs(const hd_advection *op);
/* copy phi device -> host (device_to_host != 0) or host -> device
 * (device-synchronising); the length is hd_vp_x_n_dofs */
int hd_vp_potential(hd_advection *op, double *phi, int device_to_host);
/* sum (d_e phi)^2 JxW per x component (device-synchronising) */
int hd_vp_field_energy(hd_advection *op, double *out);
/* mass, ||f||_L2, kinetic energy, momentum[3] of f (sync) */
int hd_vp_moments(hd_matrix_free *mf, const double *f, double *out,
                  void *stream);

/*
 * Learned initial guess of the potential solve.  No counterpart in the
 * reference (its CG starts from the previous phi, vlasov_poisson/include/
 * application.h:575-576); for callers that solve many Poisson problems along
 * one trajectory.  A convolutional network (four 3x3 periodic convolutions,
 * 1 -> 32 -> 32 -> 32 -> 1 channels, leaky ReLU 1/8, BF16 weights, FP32
 * accumulation and Adam state) maps the residual of the previous phi,
 * b - A phi_prev, as a normalised image of the x nodes (x cells x nodes per
 * direction, a multiple of 16 each; dim_x == 2, the 2D tensor-product
 * solver: k <= 3, at most 1024 x cells) to the shape of the correction CG
 * will find.  It trains on the operator's own converged solves, pairs
 * (b - A phi_prev, phi - phi_prev) kept in a ring of `capacity`, with
 * train_steps Adam steps on batches of train_batch every train_every solves
 * once min_pairs are stored (train_every 0: no training).  CG starts from
 * the energy-minimising combination of phi_prev and the guess and still runs
 * to poisson_tol, so the guess changes the iteration count, not the
 * result's tolerance.  The training runs on a stream the operator owns
 * (default priority, non-blocking), beside the solve's CG, ordered against
 * the caller's stream by events.  Enabling it switches the exact start off.
 */
int hd_vp_learned_guess_enable(hd_advection *op, unsigned long long seed,
                               int capacity, int train_every, int train_batch,
                               int train_steps, float learning_rate,
                               int min_pairs);
/* The training schedule (no counterpart in the reference): every one of
 * the first dense_solves solves counted from enabling trains, the later ones
 * every train_every solves (0: the network stays as trained). */
int hd_vp_learned_guess_schedule(hd_advection *op, long long dense_solves,
                                 int train_every);
int hd_vp_learned_guess_disable(hd_advection *op);
/* Adam steps taken, solves that used the guess, pairs stored in total,
 * last training loss (sync) */
int hd_vp_learned_guess_stats(hd_advection *op, long long *train_steps,
                              long long *guessed_solves, long long *pairs,
                              double *last_loss);

/*
 * The network on its own (no counterpart in the reference): images are
 * device arrays of batch x height x width FP32, 16-byte aligned; height and
 * width multiples of 16.  Parameters in the canonical FP32 layout, per layer
 * w[out][in][3][3] then b[out] (hd_cnn_n_params() floats); setting them
 * restarts Adam.  hd_cnn_loss_grad returns loss = mean (net(in) - target)^2
 * and its gradient (host arrays, either may be NULL; sync).
 */
typedef struct hd_cnn hd_cnn;
int hd_cnn_n_params(void);
int hd_cnn_create(int height, int width, unsigned long long seed,
                  hd_cnn **out);
int hd_cnn_destroy(hd_cnn *net);
int hd_cnn_params(hd_cnn *net, float *host, int device_to_host);
int hd_cnn_forward(hd_cnn *net, const float *in, float *out, int batch,
                   void *stream);
int hd_cnn_loss_grad(hd_cnn *net, const float *in, const float *target,
                     int batch, double *loss, float *grad, void *stream);
int hd_cnn_train_step(hd_cnn *net, const float *in, const float *target,
                      int batch, float learning_rate, void *stream);
int hd_cnn_last_loss(hd_cnn *net, double *loss);
/* The learned guess's conversions between a 2D x-space nodal vector
 * (n0 x n1 cells, N x N Gauss-Lobatto nodes each, cell-major) and the
 * (n1 N) x (n0 N) image the network sees (device arrays; no counterpart in
 * the reference):
 *   dofs_to_image: v = u - sub (sub may be NULL); image = (v - mean v) / rms,
 *                  *scale = rms (scale: one device double)
 *   image_to_dofs: u = scale * image, continuous across cell faces (a node
 *                  shared by neighbouring cells gets the mean of its two or
 *                  four pixels) */
int hd_cnn_dofs_to_image(const double *u, const double *sub, int n0, int n1,
                         int N, float *image, double *scale, void *stream);
int hd_cnn_image_to_dofs(const float *image, const double *scale, int n0,
                         int n1, int N, double *u, void *stream);
/* the network of an operator's learned guess (borrowed; NULL if off).  It
 * is freed by hd_vp_learned_guess_disable, by a second
 * hd_vp_learned_guess_enable and by hd_vp_exact_start(op, 1): the pointer
 * must not be used after any of them. */
hd_cnn *hd_vp_learned_guess_net(hd_advection *op);

/* dst (FP64) = function at the Gauss-Lobatto nodes (numerics/
 * vector_tools.h:81-135); params: D values or NULL. */
int hd_interpolate(hd_matrix_free *mf, int function, const double *params,
                   double time, double *dst, void *stream);

/* dst += amplitude * U(-1, 1), counter-based SplitMix64 keyed by the DoF
 * index: the seeded noise of the benchmark's initial data. */
int hd_add_noise(hd_matrix_free *mf, double *dst, uint64_t seed,
                 double amplitude, void *stream);

/* out[0] = ||u||_L2, out[1] = ||u - f(t)||_L2 at the quadrature points
 * (numerics/vector_tools.h:144-223); FP64 accumulation, synchronous. */
int hd_norm_and_error(hd_matrix_free *mf, const double *src, int function,
                      const double *params, double time, double *out,
                      void *stream);

/* hd_norm_and_error for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  It calls no allocating routine and does
 * not synchronise: out_dev is a caller-owned device array of two doubles
 * that receives (||u||_L2, ||u - f(t)||_L2), stream ordered.  Its scratch
 * (the norm kernel's block partial sums) is a fixed device array of
 * 2 x HD_NOALLOC_NORM_BLOCKS doubles (32 KiB) per GPU, held until the
 * process ends; the kernel's blocks stride over the cells, so meshes of any
 * size fit.  Cells too large for the kernel's LDS are rejected
 * (HD_ERR_UNSUPPORTED), as by hd_norm_and_error.  Calls share their GPU's
 * scratch: calls that overlap on one GPU (different streams, or threads)
 * race and must be serialised by the caller.  The first call on a device
 * looks up the scratch's device address (make it outside a graph
 * capture). */
#define HD_NOALLOC_NORM_BLOCKS 2048
int hd_norm_and_error_noalloc(hd_matrix_free *mf, const double *src,
                              int function, const double *params, double time,
                              double *out_dev, void *stream);

/* dst (precision) = src (FP64), element-wise conversion / copy. */
int hd_vector_convert(hd_matrix_free *mf, const double *src, void *dst,
                      int precision, void *stream);

/* CFL time step (operators/advection/cfl.h:94-124). */
double hd_critical_time_step(const hd_matrix_free *mf, const double *velocity);

/* Device memory helpers for hosts without another allocator; the copies are
 * device-synchronising. */
int hd_device_malloc(void **ptr, uint64_t bytes);
int hd_device_free(void *ptr);
int hd_memcpy_htod(void *dst, const void *src, uint64_t bytes);
int hd_memcpy_dtoh(void *dst, const void *src, uint64_t bytes);
int hd_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif
