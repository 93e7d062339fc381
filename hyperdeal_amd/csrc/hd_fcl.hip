// hd_fcl.hip -- the face-centric (FCL) form of the phase-space DG advection
// operator on gfx950: hyper.deal's MatrixFree::loop with the cell, face and
// boundary operations followed by the inverse mass matrix in a cell_loop
// (matrix_free/matrix_free.templates.h:1417-1465, 1634-1746;
// operators/advection/advection_operation.h:159-183, 573-1131).
//
// The vector being assembled (asm) holds, per cell, the integrals of the
// operator against the ND^D nodal test functions; three kernel kinds build
// and finish it, in stream order:
//   cell kernel   asm  = S^T^D [ cell integrand ](S^D src)         (639-818)
//   face kernels  asm += both sides' face integrals of one colour of faces
//                 of one direction, each face evaluated once          (825-1000)
//                 (boundary faces of non-periodic directions: u+ = -u- + 2g,
//                 only the inside cell is written)                   (1009-1131)
//   mass kernel   dst  = M^{-1} asm, M with the Gauss(k+1) rule of the
//                 reference's inverse-mass evaluator (FEEvaluationInverse:
//                 S^{-T}, / JxW, S^{-1}; collocation: / JxW)         (573-632)
// A face writes into the two cells it separates, so the faces of direction d
// are launched in colours: face positions p (between cells p-1 and p along
// d) of one parity never share a cell.  With periodic wrap and an odd cell
// count the wrap face p = 0 gets a colour of its own.  Within one face the
// minus cell receives at its last node layer along d and the plus cell at
// its first, so even a periodic direction of one cell is race free.
//
// Faces use the nodal trace: with Gauss-Lobatto nodes the trace of a cell's
// polynomial on a face is the polynomial through its face nodes, which S^{D-1}
// takes to the face quadrature points (FEFaceEvaluation::read_dof_values /
// evaluate, fe_evaluation_face.h:241-288; integrate is the transpose).
#include "hd_device.h"
#include "hd_functions.h"
#include "hd_quad_common.h"

namespace hd
{
namespace fcl
{
using quad::kBlock;
using quad::QTab;

// the 1D tables in the arithmetic type (FP32: converted once per block)
template <class T>
__device__ inline QTab<T> tables(const QuadDev &q, T *tabs)
{
  if constexpr (sizeof(T) == sizeof(double))
    {
      (void)tabs;
      return {q.S, q.Dq, q.F, q.Sinv};
    }
  else
    {
      for (int i = threadIdx.x; i < 64; i += blockDim.x)
        {
          tabs[i] = (T)q.S[i];
          tabs[64 + i] = (T)q.Dq[i];
          tabs[144 + i] = (T)q.Sinv[i];
          if (i < 16)
            tabs[128 + i] = (T)q.F[i];
        }
      __syncthreads();
      return {tabs, tabs + 64, tabs + 128, tabs + 144};
    }
}

// ---------------------------------------------------------------------------
// cell kernel: the cell integral tested with the nodal basis, written into
// asm (advection_operation.h:639-818)
template <int D, int ND, int NQ, class T>
__global__ __launch_bounds__(kBlock) void cell_kernel(
  const T *__restrict__ src, T *__restrict__ asmv, QuadDev q, MeshDev mesh,
  int collocation, DeformDev geo)
{
  using Sh = quad::Shape<D, ND, NQ>;
  constexpr int CELL = Sh::CELL, CELLD = Sh::CELLD, FACE = Sh::FACE;
  constexpr int CPB = Sh::CPB;
  constexpr int CELLP = CELL + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *U = reinterpret_cast<T *>(smem_raw); // CPB x CELLP
  T *R = U + CPB * CELLP;                 // CPB x CELLP
  __shared__ T tabs[64 + 64 + 16 + 64];
  const QTab<T> tb = tables<T>(q, tabs);

  const long long cell0 = (long long)blockIdx.x * CPB;
  const int ncell = (int)min((long long)CPB, mesh.n_cells - cell0);
  const int tid = threadIdx.x;

  // a line along direction 0 of a cell array (fi: its index among the
  // lines): the offset of its first node in the nodal cell and whether its
  // other digits are all < ND (it holds nodal values)
  auto line0 = [&](int fi, int &nb, bool &node) {
    int rem = fi, pwn = ND;
    nb = 0;
    node = true;
#pragma unroll
    for (int j = 1; j < D; ++j)
      {
        const int dg = rem % NQ;
        rem /= NQ;
        node = node && dg < ND;
        nb += dg * pwn;
        pwn *= ND;
      }
  };

  // nodal -> quadrature values (S: NQ x ND); the load does direction 0
  // itself (a thread reads a line of ND nodal values, contiguous in the
  // cell, and writes its NQ points; lines without nodal values are zero)
  if (!collocation)
    {
      for (int l = tid; l < ncell * FACE; l += kBlock)
        {
          const int lc = l / FACE, fi = l - lc * FACE;
          int nb;
          bool node;
          line0(fi, nb, node);
          T in[ND];
#pragma unroll
          for (int t = 0; t < ND; ++t)
            in[t] = node ? src[(cell0 + lc) * CELLD + nb + t] : T(0);
          T *pl = U + lc * CELLP + fi * NQ;
#pragma unroll
          for (int i = 0; i < NQ; ++i)
            {
              T acc = 0;
#pragma unroll
              for (int t = 0; t < ND; ++t)
                acc += tb.S[i * ND + t] * in[t];
              pl[i] = acc;
            }
        }
      __syncthreads();
      for (int d = 1; d < D; ++d)
        {
          quad::contract_lines<D, NQ, ND, NQ, T>(U, CELLP, d, tb.S, ND, 1,
                                                 ncell);
          __syncthreads();
        }
    }
  else
    {
      // collocation: the nodes are the points
      for (int i = tid; i < ncell * CELLD; i += kBlock)
        {
          const int lc = i / CELLD;
          U[lc * CELLP + (i - lc * CELLD)] = src[cell0 * CELLD + i];
        }
      __syncthreads();
    }

  // the integrand: -skew (a . grad u) JxW tested with the values plus
  // (1 - skew) u a_d JxW tested with the gradients (D_d^T)
  for (int d = 0; d < D; ++d)
    {
      const int sd = ipow(NQ, d);
      const T ad = (T)(q.a[d] / mesh.h[d]);
      const T skew = (T)q.skew;
      for (int l = tid; l < ncell * FACE; l += kBlock)
        {
          const int lc = l / FACE;
          const int base = quad::line_base<D, NQ>(l - lc * FACE, d);
          const T *u = U + lc * CELLP + base;
          T *r = R + lc * CELLP + base;
          T uq[NQ], jw[NQ], acc[NQ];
#pragma unroll
          for (int j = 0; j < NQ; ++j)
            {
              uq[j] = u[j * sd];
              // deformed: the contravariant velocity (J^{-1} a)_d times JxW
              jw[j] = geo.on ? (T)quad::deformed_velocity_jxw<NQ>(
                                 geo, cell0 + lc, base + j * sd, d)
                             : (T)quad::jxw_point<D, NQ>(q, mesh,
                                                         base + j * sd) *
                                 ad;
              acc[j] = 0;
            }
          if (q.skew != 0.0)
#pragma unroll
            for (int i = 0; i < NQ; ++i)
              {
                T g = 0;
#pragma unroll
                for (int j = 0; j < NQ; ++j)
                  g += tb.Dq[i * NQ + j] * uq[j];
                acc[i] += -skew * g * jw[i];
              }
          if (q.skew != 1.0)
#pragma unroll
            for (int i = 0; i < NQ; ++i)
              {
                T s = 0;
#pragma unroll
                for (int j = 0; j < NQ; ++j)
                  s += tb.Dq[j * NQ + i] * ((1 - skew) * uq[j] * jw[j]);
                acc[i] += s;
              }
          // direction 0 writes R, the others add
#pragma unroll
          for (int i = 0; i < NQ; ++i)
            r[i * sd] = d == 0 ? acc[i] : r[i * sd] + acc[i];
        }
      __syncthreads();
    }

  // test with the nodal basis: S^T along every direction (NQ -> ND), the
  // last one (direction 0) with the store of the line's ND nodal values
  if (!collocation)
    {
      for (int d = D - 1; d >= 1; --d)
        {
          quad::contract_lines<D, NQ, NQ, ND, T>(R, CELLP, d, tb.S, 1, ND,
                                                 ncell);
          __syncthreads();
        }
      for (int l = tid; l < ncell * FACE; l += kBlock)
        {
          const int lc = l / FACE, fi = l - lc * FACE;
          int nb;
          bool node;
          line0(fi, nb, node);
          if (!node)
            continue;
          const T *pl = R + lc * CELLP + fi * NQ;
          T v[NQ];
#pragma unroll
          for (int j = 0; j < NQ; ++j)
            v[j] = pl[j];
#pragma unroll
          for (int i = 0; i < ND; ++i)
            {
              T acc = 0;
#pragma unroll
              for (int j = 0; j < NQ; ++j)
                acc += tb.S[j * ND + i] * v[j];
              asmv[(cell0 + lc) * CELLD + nb + i] = acc;
            }
        }
    }
  else
    for (int i = tid; i < ncell * CELLD; i += kBlock)
      {
        const int lc = i / CELLD;
        asmv[cell0 * CELLD + i] = R[lc * CELLP + (i - lc * CELLD)];
      }
}

// ---------------------------------------------------------------------------
// face kernel: the faces of direction d at positions p = start + step j,
// j < count (one colour), each evaluated once and integrated into both cells
// (advection_operation.h:825-1131).  A face of the face array is the cell
// array with direction d of extent 1: NQ^{D-1} points, the nodal values at
// the points whose digits are all < ND.

// in-place 1D contraction of the lines along face direction jj (0..D-2) of
// nf face arrays: line[0, NO) <- M line[0, NI)
template <int D, int NQ, int NI, int NO, class T>
__device__ inline void contract_face_lines(T *buf, int jj, const T *M, int sr,
                                           int sc, int nf)
{
  constexpr int FACE = ipow(NQ, D - 1);
  constexpr int NL = FACE / NQ;
  const int sj = ipow(NQ, jj);
  for (int l = threadIdx.x; l < nf * NL; l += kBlock)
    {
      const int lf = l / NL;
      int li = l - lf * NL;
      int base = 0, pw = 1;
#pragma unroll
      for (int k = 0; k < D - 1; ++k)
        {
          if (k != jj)
            {
              base += (li % NQ) * pw;
              li /= NQ;
            }
          pw *= NQ;
        }
      T *p = buf + lf * FACE + base;
      T in[NI];
#pragma unroll
      for (int t = 0; t < NI; ++t)
        in[t] = p[t * sj];
#pragma unroll
      for (int i = 0; i < NO; ++i)
        {
          T acc = 0;
#pragma unroll
          for (int t = 0; t < NI; ++t)
            acc += M[i * sr + t * sc] * in[t];
          p[i * sj] = acc;
        }
    }
}

// the face's cells: cell coordinates of the D-1 other directions from `rest`,
// position p along d; minus = cell p-1 (left), plus = cell p (right), -1
// where the face is a boundary face
__device__ inline void face_cells(const MeshDev &m, const int *periodic, int d,
                                  int p, long long rest, long long &minus,
                                  long long &plus)
{
  long long base = 0;
#pragma unroll
  for (int j = 0; j < kMaxDim; ++j)
    {
      if (j >= m.dim || j == d)
        continue;
      const long long c = rest % m.n[j];
      rest /= m.n[j];
      base += c * m.stride[j];
    }
  const int n = m.n[d];
  const long long s = m.stride[d];
  if (periodic[d])
    {
      minus = base + (long long)((p + n - 1) % n) * s;
      plus = base + (long long)(p % n) * s;
    }
  else
    {
      minus = p > 0 ? base + (long long)(p - 1) * s : -1;
      plus = p < n ? base + (long long)p * s : -1;
    }
}

template <int D, int ND, int NQ, class T>
struct FaceShape
{
  static constexpr int FACE = ipow(NQ, D - 1);   // face quadrature points
  // faces per block: about 1024 face points (four per thread), at most one
  // face per thread (the face set-up uses thread f for face f)
  static constexpr int FPB = FACE >= 2048 ? 1
                             : 2048 / FACE > kBlock ? kBlock
                                                    : 2048 / FACE;
};

template <int D, int ND, int NQ, class T>
__global__ __launch_bounds__(kBlock) void face_kernel(
  const T *__restrict__ src, T *__restrict__ asmv, QuadDev q, MeshDev mesh,
  int d, int start, int step, int count, long long n_faces, double time,
  int dim_x, int collocation, DeformDev geo)
{
  using FS = FaceShape<D, ND, NQ, T>;
  constexpr int FACE = FS::FACE, FPB = FS::FPB;
  constexpr int CELLD = ipow(ND, D);
  __shared__ T Fm[FPB * FACE], Fp[FPB * FACE];
  __shared__ long long cm[FPB], cp[FPB];
  __shared__ int pos[FPB];
  __shared__ T tabs[64 + 64 + 16 + 64];
  const QTab<T> tb = tables<T>(q, tabs);

  const long long f0 = (long long)blockIdx.x * FPB;
  const int nf = (int)min((long long)FPB, n_faces - f0);
  const int tid = threadIdx.x;
  if (tid < nf)
    {
      const long long f = f0 + tid;
      const int p = start + step * (int)(f % count);
      long long mi, pl;
      face_cells(mesh, q.periodic, d, p, f / count, mi, pl);
      cm[tid] = mi;
      cp[tid] = pl;
      pos[tid] = p;
    }
  __syncthreads(); // the face set-up is visible

  // the face nodes of both sides at the face points (S along the D-1 face
  // directions; a boundary face: the missing side is 0).  The gather does
  // the first face direction j0 itself: a thread owns a line along j0 of
  // one side's face array, reads the line's ND nodes from the cell (zero
  // for lines through positions with a digit >= ND in another face
  // direction) and writes its NQ points; the passes after it the others.
  constexpr int NL = FACE / NQ;
  const int j0 = d == 0 ? 1 : 0;
  const int sn0 = ipow(ND, j0); // node stride of j0 in a cell
  const int sdn = ipow(ND, d);
  // a line of one side: its base in the face array, the node offset of its
  // first node in the side's cell (layer ND-1 of the minus cell, 0 of the
  // plus cell) and whether its other digits are all < ND
  auto face_line = [&](int ls, int side, int &fb, int &off, bool &node) {
    int rem = ls, nb = 0, pwq = 1, pwn = 1;
    fb = 0;
    node = true;
#pragma unroll
    for (int j = 0; j < D; ++j)
      {
        if (j != d)
          {
            if (j != j0)
              {
                const int dg = rem % NQ;
                rem /= NQ;
                node = node && dg < ND;
                fb += dg * pwq;
                nb += dg * pwn;
              }
            pwq *= NQ;
          }
        pwn *= ND;
      }
    off = nb + (side ? 0 : (ND - 1) * sdn);
  };
  for (int l = tid; l < 2 * nf * NL; l += kBlock)
    {
      const int side = l >= nf * NL; // 0: minus cell (Fm), 1: plus (Fp)
      const int ls = l - side * nf * NL;
      const int lf = ls / NL;
      int fb, off;
      bool node;
      face_line(ls - lf * NL, side, fb, off, node);
      const long long cl = side ? cp[lf] : cm[lf];
      T in[ND];
#pragma unroll
      for (int t = 0; t < ND; ++t)
        in[t] = node && cl >= 0 ? src[cl * CELLD + off + t * sn0] : T(0);
      T *pl = (side ? Fp : Fm) + lf * FACE + fb;
      if (collocation)
        {
#pragma unroll
          for (int t = 0; t < ND; ++t)
            pl[t] = in[t];
        }
      else
#pragma unroll
        for (int i = 0; i < NQ; ++i)
          {
            T acc = 0;
#pragma unroll
            for (int t = 0; t < ND; ++t)
              acc += tb.S[i * ND + t] * in[t];
            pl[i] = acc;
          }
    }
  __syncthreads();
  if (!collocation)
    for (int jj = 1; jj < D - 1; ++jj)
      {
        contract_face_lines<D, NQ, ND, NQ, T>(Fm, jj, tb.S, ND, 1, nf);
        contract_face_lines<D, NQ, ND, NQ, T>(Fp, jj, tb.S, ND, 1, nf);
        __syncthreads();
      }

  // the flux at the face points, submitted with the face JxW: the minus
  // side gets -(flux - skew u- an) JxW_f and the plus side the same with its
  // own normal (-an); a boundary face only its inside cell
  const double a = q.a[d];
  for (int i = tid; i < nf * FACE; i += kBlock)
    {
      const int lf = i / FACE, fi = i - lf * FACE;
      double jf = 1.0;
      {
        int rem = fi;
#pragma unroll
        for (int j = 0; j < D; ++j)
          {
            if (j == d)
              continue;
            jf *= mesh.h[j] * q.w[rem % NQ];
            rem /= NQ;
          }
      }
      const bool has_m = cm[lf] >= 0, has_p = cp[lf] >= 0;
      // deformed: (a . n) JxW_face of the cell whose face it is, from the
      // tables (the face JxW is in it: jf = 1)
      const int base = quad::line_base<D, NQ>(fi, d);
      if (has_m && has_p)
        {
          const double um = Fm[i], up = Fp[i];
          double an = a; // normal of the minus side: +e_d
          if (geo.on)
            {
              an = quad::deformed_face<NQ>(geo, cm[lf], base, d, 1);
              jf = 1.0;
            }
          const double flux = 0.5 * ((um + up) * an + fabs(an) * (um - up));
          Fm[i] = (T)(-(flux - q.skew * um * an) * jf);
          Fp[i] = (T)(-(-flux + q.skew * up * an) * jf);
        }
      else
        {
          // boundary: side 1 (right end) for the minus cell, side 0 for the
          // plus cell
          const int side = has_m ? 1 : 0;
          const double um = has_m ? (double)Fm[i] : (double)Fp[i];
          double g = 0.0;
          if (q.bc_inhomogeneous)
            {
              const long long cell = has_m ? cm[lf] : cp[lf];
              double x[kMaxDim];
              long long cc = cell;
              int rem = fi;
              for (int j = 0; j < D; ++j)
                {
                  const int cj = (int)(cc % mesh.n[j]);
                  cc /= mesh.n[j];
                  double xi;
                  if (j == d)
                    xi = (double)side;
                  else
                    {
                      xi = q.qp[rem % NQ];
                      rem /= NQ;
                    }
                  x[j] = mesh.lo[j] + (cj + xi) * mesh.h[j];
                }
              if (geo.on)
                deform_point(geo, mesh, x);
              g = eval_function(q.bc_function, q.bc_params, dim_x, D, x, time);
            }
          const double up = -um + 2.0 * g;
          double an = side ? a : -a;
          if (geo.on)
            {
              an = quad::deformed_face<NQ>(geo, has_m ? cm[lf] : cp[lf], base,
                                           d, side);
              jf = 1.0;
            }
          const double flux = 0.5 * ((um + up) * an + fabs(an) * (um - up));
          const T v = (T)(-(flux - q.skew * um * an) * jf);
          if (has_m)
            Fm[i] = v;
          else
            Fp[i] = v;
        }
    }
  __syncthreads();
  // test with the face's nodal basis (S^T along the face directions), the
  // first face direction j0 together with the add into the cells: the
  // minus cell's last node layer along d, the plus cell's first
  if (!collocation)
    for (int jj = D - 2; jj >= 1; --jj)
      {
        contract_face_lines<D, NQ, NQ, ND, T>(Fm, jj, tb.S, 1, ND, nf);
        contract_face_lines<D, NQ, NQ, ND, T>(Fp, jj, tb.S, 1, ND, nf);
        __syncthreads();
      }
  for (int l = tid; l < 2 * nf * NL; l += kBlock)
    {
      const int side = l >= nf * NL;
      const int ls = l - side * nf * NL;
      const int lf = ls / NL;
      int fb, off;
      bool node;
      face_line(ls - lf * NL, side, fb, off, node);
      const long long cl = side ? cp[lf] : cm[lf];
      if (!node || cl < 0)
        continue;
      const T *pl = (side ? Fp : Fm) + lf * FACE + fb;
      T v[NQ];
#pragma unroll
      for (int j = 0; j < NQ; ++j)
        v[j] = pl[j];
      T *out = asmv + cl * CELLD + off;
#pragma unroll
      for (int i = 0; i < ND; ++i)
        {
          T acc;
          if (collocation)
            acc = v[i];
          else
            {
              acc = 0;
#pragma unroll
              for (int j = 0; j < NQ; ++j)
                acc += tb.S[j * ND + i] * v[j];
            }
          out[i * sn0] += acc;
        }
    }
}

// ---------------------------------------------------------------------------
// inverse mass kernel (advection_operation.h:573-632): dst = M^{-1} asm with
// M^{-1} = S^{-1} (JxW)^{-1} S^{-T} of the Gauss(k+1) rule (collocation: the
// Gauss-Lobatto rule, S = I).  MODE 0: write dst; MODE 1: LSRK stage
// epilogue (sol += b_dt K, ti_out = sol_old + a_dt K; asm may be ti_out).
// asmv may be dst (apply) or ti_out (stage): a block reads its cells into
// LDS before it writes them.
// cells per block of the inverse mass kernel: about 2048 nodal values
__host__ __device__ constexpr int mass_cpb(int celld)
{
  return celld >= 2048 ? 1 : 2048 / celld;
}

template <int D, int ND, int MODE, class T>
__global__ __launch_bounds__(kBlock) void mass_kernel(const T *asmv, T *dst,
                                                      MassDev md, MeshDev mesh,
                                                      StageDev stage,
                                                      DeformDev geo)
{
  constexpr int CELLD = ipow(ND, D);
  constexpr int CPB = mass_cpb(CELLD);
  constexpr int CELLP = CELLD + 1;
  __shared__ T U[CPB * CELLP];
  __shared__ T sinv[64];
  for (int i = threadIdx.x; i < ND * ND; i += kBlock)
    sinv[i] = (T)md.Sinv[i];
  const long long cell0 = (long long)blockIdx.x * CPB;
  const int ncell = (int)min((long long)CPB, mesh.n_cells - cell0);
  const int tid = threadIdx.x;
  constexpr int NL = CELLD / ND; // lines along a direction per cell
  if (!md.collocation)
    {
      __syncthreads(); // the table is visible
      // Sinv^T along direction 0 with the load (a thread reads a line of
      // ND contiguous values), then the other directions; A(r, c) =
      // Sinv[c][r]
      for (int l = tid; l < ncell * NL; l += kBlock)
        {
          const int lc = l / NL, fi = l - lc * NL;
          T in[ND];
#pragma unroll
          for (int t = 0; t < ND; ++t)
            in[t] = asmv[(cell0 + lc) * CELLD + fi * ND + t];
          T *pl = U + lc * CELLP + fi * ND;
#pragma unroll
          for (int i = 0; i < ND; ++i)
            {
              T acc = 0;
#pragma unroll
              for (int t = 0; t < ND; ++t)
                acc += sinv[t * ND + i] * in[t];
              pl[i] = acc;
            }
        }
      __syncthreads();
      for (int d = 1; d < D; ++d)
        {
          quad::contract_lines<D, ND, ND, ND, T>(U, CELLP, d, sinv, 1, ND,
                                                 ncell);
          __syncthreads();
        }
    }
  else
    {
      for (int i = tid; i < ncell * CELLD; i += kBlock)
        {
          const int lc = i / CELLD;
          U[lc * CELLP + (i - lc * CELLD)] = asmv[cell0 * CELLD + i];
        }
      __syncthreads();
    }
  for (int i = tid; i < ncell * CELLD; i += kBlock)
    {
      const int lc = i / CELLD;
      int li = i - lc * CELLD;
      double jxw = 1.0;
      if (geo.on) // the mass rule is the operator's (ND points)
        jxw = quad::deformed_jxw<ND>(geo, cell0 + lc, li);
      else
#pragma unroll
        for (int j = 0; j < D; ++j)
          {
            jxw *= mesh.h[j] * md.w[li % ND];
            li /= ND;
          }
      U[lc * CELLP + (i - lc * CELLD)] /= (T)jxw;
    }
  __syncthreads();
  auto store = [&](long long g, T k) {
    if constexpr (MODE == 0)
      dst[g] = k;
    else
      {
        const double s = stage.sol[g];
        stage.sol[g] = s + stage.b_dt * (double)k;
        if (stage.ti_out)
          static_cast<T *>(stage.ti_out)[g] = (T)(s + stage.a_dt * (double)k);
      }
  };
  if (!md.collocation)
    {
      for (int d = D - 1; d >= 1; --d)
        {
          quad::contract_lines<D, ND, ND, ND, T>(U, CELLP, d, sinv, ND, 1,
                                                 ncell);
          __syncthreads();
        }
      // Sinv along direction 0 with the store of the line's values
      for (int l = tid; l < ncell * NL; l += kBlock)
        {
          const int lc = l / NL, fi = l - lc * NL;
          const T *pl = U + lc * CELLP + fi * ND;
          T v[ND];
#pragma unroll
          for (int t = 0; t < ND; ++t)
            v[t] = pl[t];
#pragma unroll
          for (int i = 0; i < ND; ++i)
            {
              T acc = 0;
#pragma unroll
              for (int t = 0; t < ND; ++t)
                acc += sinv[i * ND + t] * v[t];
              store((cell0 + lc) * CELLD + fi * ND + i, acc);
            }
        }
    }
  else
    for (int i = tid; i < ncell * CELLD; i += kBlock)
      {
        const int lc = i / CELLD;
        store(cell0 * CELLD + i, U[lc * CELLP + (i - lc * CELLD)]);
      }
}

// the colours of direction d: (start, step, count) of face positions
struct Colour
{
  int start, step, count;
};

inline int colours(int n, bool periodic, Colour *c)
{
  if (periodic && n % 2 == 1)
    {
      if (n == 1)
        {
          c[0] = {0, 1, 1};
          return 1;
        }
      c[0] = {2, 2, (n - 1) / 2}; // 2, 4, ..., n-1
      c[1] = {1, 2, (n - 1) / 2}; // 1, 3, ..., n-2
      c[2] = {0, 1, 1};           // the wrap face
      return 3;
    }
  const int np = periodic ? n : n + 1;
  c[0] = {0, 2, (np + 1) / 2};
  c[1] = {1, 2, np / 2};
  return c[1].count > 0 ? 2 : 1;
}

template <int D, int ND, int NQ, class T>
hipError_t launch(int mode, const T *src, T *asmv, T *dst, const QuadDev &q,
                  const MassDev &md, const MeshDev &mesh,
                  const StageDev &stage, double time, int dim_x,
                  int collocation, const DeformDev &geo, hipStream_t stream)
{
  using Sh = quad::Shape<D, ND, NQ>;
  // 1) cells
  {
    const size_t lds = sizeof(T) * 2 * Sh::CPB * (Sh::CELL + 1);
    hipError_t e = hipFuncSetAttribute(
      reinterpret_cast<const void *>(&cell_kernel<D, ND, NQ, T>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
      return e;
    const long long grid = (mesh.n_cells + Sh::CPB - 1) / Sh::CPB;
    hipLaunchKernelGGL((cell_kernel<D, ND, NQ, T>), dim3((unsigned)grid),
                       dim3(kBlock), lds, stream, src, asmv, q, mesh,
                       collocation, geo);
    e = hipGetLastError();
    if (e != hipSuccess)
      return e;
  }
  // 2) faces, direction by direction, colour by colour
  using FS = FaceShape<D, ND, NQ, T>;
  for (int d = 0; d < D; ++d)
    {
      Colour c[3];
      const int nc = colours(mesh.n[d], q.periodic[d] != 0, c);
      const long long others = mesh.n_cells / mesh.n[d];
      for (int k = 0; k < nc; ++k)
        {
          if (c[k].count == 0)
            continue;
          const long long n_faces = others * c[k].count;
          const long long grid = (n_faces + FS::FPB - 1) / FS::FPB;
          hipLaunchKernelGGL((face_kernel<D, ND, NQ, T>), dim3((unsigned)grid),
                             dim3(kBlock), 0, stream, src, asmv, q, mesh, d,
                             c[k].start, c[k].step, c[k].count, n_faces, time,
                             dim_x, collocation, geo);
          const hipError_t e = hipGetLastError();
          if (e != hipSuccess)
            return e;
        }
    }
  // 3) inverse mass matrix
  constexpr int CPB = mass_cpb(ipow(ND, D));
  const long long grid = (mesh.n_cells + CPB - 1) / CPB;
  if (mode == 0)
    hipLaunchKernelGGL((mass_kernel<D, ND, 0, T>), dim3((unsigned)grid),
                       dim3(kBlock), 0, stream, asmv, dst, md, mesh, stage,
                       geo);
  else
    hipLaunchKernelGGL((mass_kernel<D, ND, 1, T>), dim3((unsigned)grid),
                       dim3(kBlock), 0, stream, asmv, dst, md, mesh, stage,
                       geo);
  return hipGetLastError();
}

template <class T>
hipError_t dispatch(int D, int ND, int NQ, int mode, const T *src, T *asmv,
                    T *dst, const QuadDev &q, const MassDev &md,
                    const MeshDev &mesh, const StageDev &stage, double time,
                    int dim_x, int collocation, const DeformDev &geo,
                    hipStream_t stream)
{
#define HD_CASE(DD, ND_, NQ_)                                                \
  if (D == DD && ND == ND_ && NQ == NQ_)                                     \
    return launch<DD, ND_, NQ_, T>(mode, src, asmv, dst, q, md, mesh, stage, \
                                   time, dim_x, collocation, geo, stream);
#define HD_CASE2(DD, ND_) HD_CASE(DD, ND_, ND_) HD_CASE(DD, ND_, ND_ + 1)
  HD_CASE2(2, 2) HD_CASE2(2, 3) HD_CASE2(2, 4) HD_CASE2(2, 5) HD_CASE(2, 6, 6)
  HD_CASE2(3, 2) HD_CASE2(3, 3) HD_CASE2(3, 4) HD_CASE2(3, 5) HD_CASE(3, 6, 6)
  HD_CASE2(4, 2) HD_CASE2(4, 3) HD_CASE2(4, 4) HD_CASE2(4, 5) HD_CASE(4, 6, 6)
  HD_CASE2(5, 2) HD_CASE2(5, 3) HD_CASE2(5, 4)
  HD_CASE2(6, 2) HD_CASE2(6, 3) HD_CASE(6, 4, 4)
#undef HD_CASE2
#undef HD_CASE
  return hipErrorInvalidValue;
}
} // namespace fcl

bool fcl_supported(int D, int ND, int NQ)
{
  return quadrature_supported(D, ND, NQ);
}

hipError_t fcl_apply(int D, int ND, int NQ, int mode, const double *src,
                     double *asmv, double *dst, const QuadDev &q,
                     const MassDev &md, const MeshDev &mesh,
                     const StageDev &stage, double time, int dim_x,
                     int collocation, const DeformDev &geo,
                     hipStream_t stream)
{
  return fcl::dispatch<double>(D, ND, NQ, mode, src, asmv, dst, q, md, mesh,
                               stage, time, dim_x, collocation, geo, stream);
}

hipError_t fcl_apply_f32(int D, int ND, int NQ, int mode, const float *src,
                         float *asmv, float *dst, const QuadDev &q,
                         const MassDev &md, const MeshDev &mesh,
                         const StageDev &stage, double time, int dim_x,
                         int collocation, const DeformDev &geo,
                         hipStream_t stream)
{
  return fcl::dispatch<float>(D, ND, NQ, mode, src, asmv, dst, q, md, mesh,
                              stage, time, dim_x, collocation, geo, stream);
}
} // namespace hd
