// hd_device.h -- shared device-side definitions for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hd
{
constexpr int kMaxDim = 6;

__host__ __device__ constexpr int ipow(int b, int e)
{
  return e == 0 ? 1 : b * ipow(b, e - 1);
}

// Cartesian phase-space mesh as the kernels see it: cells numbered
// lexicographically, direction 0 fastest (x directions first, then v).
struct MeshDev
{
  int dim;
  int n[kMaxDim];           // cells per direction
  long long stride[kMaxDim]; // cell-id stride per direction
  long long n_cells;
  double lo[kMaxDim], h[kMaxDim];
  // cell id = inner * (lexicographic id) + sub-cell index: > 1 only for the
  // slab view of the split kernels (a D-dimensional cell seen as inner
  // cells of its first four directions)
  int inner = 1;
};

// Which neighbour feeds the inflow face of each direction (constant
// velocity): +1 -> the left neighbour's last node, -1 -> the right
// neighbour's first node, 0 -> no transport in this direction.
struct InflowDev
{
  int side[kMaxDim];
};

// LSRK stage epilogue: sol += b_dt K; ti_out = sol_old + a_dt K
struct StageDev
{
  double a_dt, b_dt;
  double *sol;   // FP64 solution, updated in place
  void *ti_out;  // next stage input (Number), may be null (last stage)
  void *scratch; // a free Number vector for the last stage, or null
  // Vlasov-Poisson stages: the kernel writes the density partials of
  // ti_out here (one per v cell and x node), or null
  double *rho_part;
};

// blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of logical blocks so that face neighbours share an L2
__device__ inline long long xcd_chunked_block(long long n_logical)
{
  const long long b = blockIdx.x;
  const long long per = (n_logical + 7) / 8;
  return (b % 8) * per + b / 8;
}

// neighbour cell across direction d on the inflow side (periodic)
__device__ inline long long inflow_neighbour(const MeshDev &m, long long cell,
                                             int d, int side)
{
  const long long s = m.stride[d];
  const int c = (int)((cell / s) % m.n[d]);
  if (side > 0) // from the left
    return c == 0 ? cell + (long long)(m.n[d] - 1) * s : cell - s;
  return c == m.n[d] - 1 ? cell - (long long)(m.n[d] - 1) * s : cell + s;
}

// constants of one quadrature-form operator: 1D tables (N x N each, N <= 8),
// velocity, skew, boundary description
struct QuadDev
{
  double S[64];     // nq x nd
  double Dq[64];    // nq x nq
  double F[16];     // 2 x nq
  double Sinv[64];  // nd x nq
  double w[8];      // quadrature weights on [0,1]
  double qp[8];     // quadrature points on [0,1]
  double a[kMaxDim];
  double skew;
  int periodic[kMaxDim];
  int bc_inhomogeneous; // Dirichlet: 1 -> u+ = -u- + 2 g, 0 -> u+ = -u-
  int bc_function;
  double bc_params[kMaxDim + 2];
  // Vlasov-Poisson velocity: x-velocity = v at the v quadrature point,
  // v-velocity = grad phi at the x quadrature point (field array
  // [x cell][x quadrature point][dim_x])
  int vp;
  long long n_x_cells;
};

// A deformed phase-space mesh (see hd_geometry.h): each space maps
// x^ -> x^ + delta prod_j sin(freq pi (x^_j - lo_j) / (hi_j - lo_j)) (1..1).
// The per-point tables are per space (x: [x cell][x point], v likewise):
//   jw: |det J| w at the cell points (mesh tables)
//   ca: operator tables [cell][point][dim + 1] = (J^{-1} a, |det J| w)
//   fg: operator tables [cell][2 dim][face point] = +-(J^{-1} a)_d |det J| w
// on = 0: the Cartesian mesh (no tables).
struct DeformDev
{
  int on;
  int dim_x, dim_v;
  double delta[2];
  int freq[2];
  long long n_x_cells;
  const double *xjw, *vjw;
  const double *xca, *vca;
  const double *xfg, *vfg;
};

// the physical point of the phase-space point x^ (in place)
__host__ __device__ inline void deform_point(const DeformDev &g,
                                             const MeshDev &m, double *x)
{
  const double pi = 3.14159265358979323846;
  for (int s = 0; s < 2; ++s)
    {
      const int b = s == 0 ? 0 : g.dim_x, e = s == 0 ? g.dim_x : g.dim_x + g.dim_v;
      double sv = g.delta[s];
      for (int j = b; j < e; ++j)
        sv *= sin(g.freq[s] * pi * (x[j] - m.lo[j]) / (m.n[j] * m.h[j]));
      for (int j = b; j < e; ++j)
        x[j] += sv;
    }
}

// the inverse mass matrix of the face-centric form: S^{-1} of the Gauss
// (k+1) rule (ND x ND) and its weights on [0,1]; collocation: the
// Gauss-Lobatto weights, no S
struct MassDev
{
  double Sinv[64];
  double w[8];
  int collocation;
};

// an analytic function and the 1D nodes it is interpolated at
struct FuncDev
{
  int id;
  double prm[kMaxDim + 2];
  double nodes[16]; // 1D nodal points on [0,1]
  int nd;
  int dim_x;
};

// tables of the L2 norm: nodal values -> quadrature points
struct NormDev
{
  double S[256]; // nq x nd (nq, nd <= 16)
  double qp[16], w[16];
  int nd, nq;
};

// Vlasov-Poisson: the v quadrature weights of the nodal v points
struct VpDev
{
  double wv[64]; // JxW_v at the Gauss-Lobatto v nodes ((k+1)^dim_v <= 64)
  int dofs_x, dofs_v;
};

// the x-space (Poisson) problem: exact 1D matrices of the nodal basis
struct XDev
{
  int dim;
  int n[3];
  double h[3];
  long long n_cells;
  double kk;      // penalty factor k (k+1)
  double M[64];   // int l_i l_j on [0,1]          (N x N)
  double K[64];   // int l_i' l_j' on [0,1]
  double d0[8];   // l_i'(0)
  double d1[8];   // l_i'(1)
  double S[64];   // l_j at the quadrature points  (N x N)
  double G[64];   // l_j' at the quadrature points
};

// Batched collapsed apply: several independent operators of one dimension
// D in one launch.
constexpr int kBatchMax = 64;

// per operator, constant over calls (device table)
struct BatchOp
{
  const void *coef; // FP64 or FP32 collapsed coefficients
  MeshDev mesh;
  InflowDev inflow;
  int N;
  long long n_dofs;
};

// a block's work item: logical units lb, lb + stride, ... below le of
// operator op
struct BatchTask
{
  int op;
  int stride;
  long long lb, le;
};

// the work split of a batch launch: 8 x per blocks (blocks are dispatched
// round-robin over the 8 XCDs); block `slot` of XCD x runs the tasks
// [block_start[x per + slot], block_start[x per + slot + 1]) in order
struct BatchPlan
{
  const int *block_start; // device array of 8 per + 1 entries
  int per;                // blocks per XCD
};

// per call: the vectors of every operator (kernel argument) and the stage
// coefficients
struct BatchPtrs
{
  const void *src[kBatchMax];
  void *dst[kBatchMax];
  double *sol[kBatchMax];
  void *ti[kBatchMax];
  double a_dt, b_dt;
};

// prefix sums of the operators' DoF counts (batched elementwise kernels)
struct BatchSizes
{
  long long off[kBatchMax + 1];
  int n;
};

// host launchers (one translation unit per kernel family)
bool collapsed_supported(int D, int N);
hipError_t collapsed_apply_f64(int D, int N, int mode, const double *src,
                               double *dst, const double *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream);
hipError_t collapsed_apply_f32(int D, int N, int mode, const float *src,
                               float *dst, const float *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream);
// 2+2 (D = 4) of low degree through the wave-per-group kernel
// (hd_cvwave.hip); batches of such operators likewise (units: groups of
// collapsed_wave_cells_per_unit cells, one wave per block)
bool collapsed_wave_supported(int D, int N, int precision);
int collapsed_wave_cells_per_unit(int N);
int collapsed_wave_blocks_per_cu(int N, int precision);
hipError_t collapsed_wave_f64(int N, int mode, const double *src, double *dst,
                              const double *coef, const MeshDev &mesh,
                              const InflowDev &inflow, const StageDev &stage,
                              hipStream_t stream);
hipError_t collapsed_wave_f32(int N, int mode, const float *src, float *dst,
                              const float *coef, const MeshDev &mesh,
                              const InflowDev &inflow, const StageDev &stage,
                              hipStream_t stream);
// batches: one kernel per degree N.  The dimensions and degrees with a
// batch kernel, the cells one block handles at a time (a task's units are
// multiples of it), the blocks per XCD of degree N's kernel, the batched
// apply / stage of the operators of degree N listed in `tasks` (mode as
// above) and the batched FP64 -> precision copy of sol into dst
bool collapsed_batch_supported(int D, int N);
int collapsed_batch_cells_per_block(int D, int N, int precision);
int collapsed_batch_blocks_per_xcd(int N, int precision);
hipError_t collapsed_batch(int D, int N, int precision, int mode,
                           const BatchOp *ops, const BatchTask *tasks,
                           const BatchPlan &plan, const BatchPtrs &ptrs,
                           hipStream_t stream);
hipError_t collapsed_wave_batch(int N, int precision, int mode,
                                const BatchOp *ops, const BatchTask *tasks,
                                const BatchPlan &plan, const BatchPtrs &ptrs,
                                hipStream_t stream);
hipError_t batch_convert(int precision, int n_ops, const long long *n_dofs,
                         const BatchPtrs &ptrs, hipStream_t stream);
// operators too large for one cell in LDS (D >= 5, N beyond the
// collapsed kernel's range): the first four directions on the slabs of
// fixed higher indices by the 4D kernel, the others on the planes of fixed
// lower indices; part: the first part's result (may be dst or ti_out)
bool collapsed_split_supported(int D, int N);
hipError_t collapsed_split_f64(int D, int N, int mode, const double *src,
                               double *dst, double *part, const double *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream);
hipError_t collapsed_split_f32(int D, int N, int mode, const float *src,
                               float *dst, float *part, const float *coef,
                               const MeshDev &mesh, const InflowDev &inflow,
                               const StageDev &stage, hipStream_t stream);
// FP32 3+3 cells larger than LDS in one pass, the cell's result held in a
// block's registers (hd_cres.hip)
bool collapsed_cres_supported(int D, int N);
hipError_t collapsed_cres_f32(int N, int mode, const float *src, float *dst,
                              const float *coef, const MeshDev &mesh,
                              const InflowDev &inflow, const StageDev &stage,
                              hipStream_t stream);
// ND = k+1 nodes, NQ quadrature points per direction (NQ = ND or ND + 1)
bool quadrature_supported(int D, int ND, int NQ);
// geo: the deformed mesh's tables (geo.on = 0: Cartesian)
// Dirichlet boundaries of the collapsed form (hd_dirichlet.hip): after a
// collapsed apply (mode 0) or stage (mode 1; ti_out null: the last stage)
// that read periodic images at the box edges, replace the inflow faces of
// the non-periodic directions by the reference's exterior value
// u+ = -u- + 2 g (q.bc_inhomogeneous) or -u-, at time `time`
bool dirichlet_collapsed_supported(int D, int NQ, bool inhomogeneous);
hipError_t dirichlet_correct_f64(int D, int N, int NQ, int mode,
                                 const double *src, double *dst,
                                 const double *coef, const MeshDev &mesh,
                                 const InflowDev &inflow, const QuadDev &q,
                                 const StageDev &stage, double time,
                                 int dim_x, hipStream_t stream);
hipError_t dirichlet_correct_f32(int D, int N, int NQ, int mode,
                                 const float *src, float *dst,
                                 const float *coef, const MeshDev &mesh,
                                 const InflowDev &inflow, const QuadDev &q,
                                 const StageDev &stage, double time,
                                 int dim_x, hipStream_t stream);
hipError_t quadrature_apply(int D, int ND, int NQ, int mode,
                            const double *src, double *dst, const QuadDev &q,
                            const MeshDev &mesh, const StageDev &stage,
                            double time, int dim_x, int collocation,
                            const double *efield, const DeformDev &geo,
                            hipStream_t stream);
// the FP32-mixed form: FP32 vectors and contractions, FP64 LSRK update
hipError_t quadrature_apply_f32(int D, int ND, int NQ, int mode,
                                const float *src, float *dst,
                                const QuadDev &q, const MeshDev &mesh,
                                const StageDev &stage, double time,
                                int dim_x, int collocation,
                                const double *efield, const DeformDev &geo,
                                hipStream_t stream);
// the face-centric form (MatrixFree::loop + the inverse-mass cell_loop):
// cell, face (per direction and colour) and inverse-mass kernels; asmv is
// the vector the integrals are assembled in (dst for an apply, ti_out or a
// free vector for a stage; never src)
bool fcl_supported(int D, int ND, int NQ);
hipError_t fcl_apply(int D, int ND, int NQ, int mode, const double *src,
                     double *asmv, double *dst, const QuadDev &q,
                     const MassDev &md, const MeshDev &mesh,
                     const StageDev &stage, double time, int dim_x,
                     int collocation, const DeformDev &geo,
                     hipStream_t stream);
hipError_t fcl_apply_f32(int D, int ND, int NQ, int mode, const float *src,
                         float *asmv, float *dst, const QuadDev &q,
                         const MassDev &md, const MeshDev &mesh,
                         const StageDev &stage, double time, int dim_x,
                         int collocation, const DeformDev &geo,
                         hipStream_t stream);
// geo (interpolation, norm): the deformed mesh's map and JxW tables
hipError_t launch_interpolate(double *dst, const MeshDev &mesh,
                              const FuncDev &f, double t, long long n_dofs,
                              int dofs_cell, const DeformDev &geo,
                              hipStream_t stream);
hipError_t launch_noise(double *dst, uint64_t seed, double amplitude,
                        long long n, hipStream_t stream);
hipError_t launch_convert_f32(const double *src, float *dst, long long n,
                              hipStream_t stream);
// rho = int f dv from f (FP64) or f32; copy32 / copy64 (may be null) get a
// copy of the values read (one pass over f for the density and the first
// LSRK stage's input)
hipError_t vp_density(const double *f, const float *f32, const VpDev &vp,
                      long long n_x_cells, long long n_v_cells,
                      double *partial, int n_chunks, double *rho,
                      hipStream_t stream, float *copy32 = nullptr,
                      double *copy64 = nullptr);
bool vp_poisson_supported(int dim_x, int N, long long n_x_cells);
// stats: [0] iterations (summed), [1] solves, [2] final relative residual,
// [3] relative residual of the start, [4] of the previous potential,
// [5] solves started from a guess, [6], [7] the weights of the previous
// potential and of the guess in the start (doubles in slots 2-4, 6, 7)
// tab2d: the 2D tensor-product SIPG table (vp_poisson2d_table, 8 N^2
// doubles on the device) selecting the fast 2D solver where it applies, or
// null.  residual (2D solver only): write r = b - A phi as the image
// (r - mean r) / scale of the x nodes, scale = rms(r - mean r) (the layout
// of cnn_dofs_to_image), and stop (phi, efield and stats untouched).  guess
// (2D solver: its nodal mean removed): CG starts from the
// energy-minimising combination of phi and the guess, or from the guess
// alone when guess_alone (an exact start).
struct ResidualImage
{
  float *image = nullptr;  // (n1 N) x (n0 N) floats
  double *scale = nullptr; // one double
};
hipError_t vp_poisson(const XDev &xd, int N, const double *rho, double *phi,
                      const double *guess, double *efield, double *work,
                      double rel_tol, int max_it, long long *stats,
                      hipStream_t stream, const double *tab2d = nullptr,
                      const ResidualImage *residual = nullptr,
                      bool guess_alone = false);
bool vp_poisson2d_supported(int dim_x, int N, long long n_x_cells);
// the 2+2 Vlasov-Poisson apply with the velocity (v, grad phi): a table of
// vp_apply_table_size(N) doubles, then the same values as floats (the
// FP32-mixed kernel's coefficients) (the unit-speed 1D operators of the four
// directions for both signs, S, S^-1, Gauss points), then, in doubles, the
// v factors S^-1 diag(v_d(q)+/-) S per v_d cell ([d][cell][sign][N][N]);
// precision 0 FP64, 1 FP32-mixed; mode 0 apply, 1 LSRK stage
bool vp_apply_supported(int dim_x, int dim_v, int N);
// rho = the sum over v cells of the partials a phase-space stage wrote
hipError_t vp_density_from_partials(const double *part, long long n_v_cells,
                                    long long n_x_dofs, double *rho,
                                    hipStream_t stream);
int vp_apply_table_size(int N);
// the 2+2 phase-space apply / LSRK stage (hd_vpwave.hip)
hipError_t vp_wave_apply(int N, int mode, int precision, const void *src,
                         void *dst, const double *tab, const MeshDev &mesh,
                         long long n_x_cells, const double *efield,
                         const StageDev &stage, hipStream_t stream);
void vp_poisson2d_table(const XDev &xd, int N, double *tab);
// The exact start of the 2D solver by fast diagonalisation (hd_fdm.hip):
// the generalised eigenvectors of the two 1D periodic SIPG operators on the
// device; vp_fdm_solve writes the solution of the system the 2D solver
// forms from rho into x (cell-major; one plus four small launches)
struct FdmDev
{
  double *mem = nullptr;
  double *V0 = nullptr, *V1 = nullptr; // (n_d N)^2 each, row-major
  double *inv = nullptr;               // 1 / (lambda1_r + lambda0_c)
  double *B = nullptr, *T = nullptr;   // work arrays, n1 N x n0 N
  double *x = nullptr;                 // the solution, cell-major
  int N = 0, n0 = 0, n1 = 0;
};
bool vp_fdm_supported(int dim_x, int N, const int *n, long long n_x_cells);
// tab_host: the table of vp_poisson2d_table
hipError_t vp_fdm_create(const double *tab_host, int N, int n0, int n1,
                         FdmDev *out);
void vp_fdm_destroy(FdmDev &f);
// tab: the same table on the device
hipError_t vp_fdm_solve(const FdmDev &f, const double *tab, const double *rho,
                        hipStream_t stream);
hipError_t vp_moments(const double *f, const MeshDev &mesh, const NormDev &nt,
                      int dim_x, int vol, double *partial, hipStream_t stream);
// blocks <= 0: one block per cell (partial: 2 per cell); otherwise `blocks`
// blocks striding over the cells (partial: 2 per block)
hipError_t launch_norm(const double *src, const MeshDev &mesh,
                       const NormDev &nt, const FuncDev &f, double t, int vol,
                       double *partial, const DeformDev &geo,
                       hipStream_t stream, int blocks = 0);
// out[0..1] = sqrt of the sums of n partial pairs (device pointers)
hipError_t launch_norm_finish(const double *partial, int n, double *out,
                              hipStream_t stream);

// SplitMix64 output for counter i (state seed + (i+1) golden increments),
// mapped to [-1, 1)
__host__ __device__ inline double splitmix_uniform(uint64_t seed, uint64_t i)
{
  uint64_t z = seed + (i + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
} // namespace hd
