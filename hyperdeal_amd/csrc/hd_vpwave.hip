// hd_vpwave.hip -- the 2+2 Vlasov-Poisson phase-space apply and fused LSRK
// stage (examples/vlasov_poisson/include/application.h:516-602 with the
// velocity field of velocity_field_view.h:28-209), one wave per group of
// cells, gfx950.
//
// The operator per cell, exact for affine Cartesian cells and N = k+1 Gauss
// points (every quadrature the reference applies is reproduced):
//   x_d (d = 0, 1): V_d+ (x) (A_d+ u + b_d+ u_left) + V_d- (x) (A_d- u +
//     b_d- u_right), A_d, b_d the unit-speed collapsed 1D operator along x_d
//     and V_d+/- = S^-1 diag(max / min(v_d(q), 0)) S along v_d (a factor of
//     the v cell: most cells need one sign);
//   v_e (e = 0, 1): in the x Gauss basis (S (x) S along x0, x1) every v_e
//     line has the constant speed E_e(q_x): E_e(q_x) (A_e+ or A_e-, upwind
//     by its sign) along v_e, then back with S^-1 (x) S^-1.
//
// Work mapping: a wave holds CPW = 64 / N^2 cells of one v cell (consecutive
// x cells), N^2 lanes per cell.  A lane is either an x-plane (the N^2
// values of one v node: x-layout) or a v-plane (the N^2 values of one x
// node or x Gauss point: v-layout); every 1D operator is applied inside a
// lane's registers, and the wave moves between the two layouts through its
// own LDS (rows of an odd stride: conflict-free in both layouts).  No
// workgroup barrier: every LDS exchange is inside one wave.
//   stage     u (coalesced global loads) -> LDS -> x-planes
//   X1        P_d = A_d u + b_d face_d (d = 0, 1)            x-layout
//   V1        nodal = V_0 P_0 + V_1 P_1                      v-layout
//   X1b       u^ = (S x S) u, the v faces likewise            x-layout
//   V2        w = sum_e E_e (A_e u^ + b_e face^_e)           v-layout
//   X2        out = (S^-1 x S^-1) w + nodal                   x-layout
//   out       coalesced stores (LSRK: sol += b dt out, ti = sol + a dt out)
//   density   int ti dv per x node, FP64 from the unrounded stage value
// Faces: the x inflow face of the v cell's sign (both only in a v cell that
// contains v_d = 0), the v faces each cell's E signs select.
#include "hd_device.h"
#include "hd_vp_table.h"
#include "hd_wave_common.h"

#include <algorithm>
#include <type_traits>

namespace hd
{
namespace vlasov
{
namespace vpw
{
using wave::cptr;

template <int N>
struct Geo
{
  static constexpr int N2 = N * N, N4 = N2 * N2;
  static constexpr int CPW = 64 / N2;               // cells per wave
  static constexpr int LANES = CPW * N2;            // busy lanes
  static constexpr int P = N2 | 1;                  // LDS row stride (odd)
  static constexpr int NP = 4 * N;                  // v-face planes per cell
  static constexpr int RP = (NP + N2 - 1) / N2;     // v-face planes per lane
  // LDS per cell: B0 (N^2 rows) and B1, whose rows hold the v-face planes
  // between X1b and V2
  static constexpr int R1 = NP > N2 ? NP : N2;
  static constexpr int SLOT = (N2 + R1) * P;
  static constexpr int KU = (CPW * N4 + 63) / 64;   // coalesced values/lane
};

using wave::along0;
using wave::along0_acc;
using wave::along1;
using wave::along1_acc;
using wave::launder;
using wave::nbr;
using wave::wave_sync;

template <int N, typename T, int MODE>
__global__ __launch_bounds__(64) void stage_kernel(
  const T *__restrict__ src, T *__restrict__ dst,
  const double *__restrict__ tab_in, MeshDev mesh, long long n_x_cells,
  const double *__restrict__ efield, StageDev stage, long long x_groups,
  long long n_groups)
{
  using TB = vptab::Tab<N>;
  using G = Geo<N>;
  constexpr int N2 = G::N2, N4 = G::N4, CPW = G::CPW, P = G::P;
  constexpr int KU = G::KU, RP = G::RP, NP = G::NP, SLOT = G::SLOT;
  const cptr<double> tab0 = (cptr<double>)tab_in;
  cptr<double> tab = tab0;
  cptr<float> tabf = (cptr<float>)(tab_in + TB::size);
  auto tc = [&](int i) -> T {
    if constexpr (sizeof(T) == 4)
      return tabf[i];
    else
      return tab[i];
  };
  const int n0 = mesh.n[0], n1 = mesh.n[1], nv0 = mesh.n[2],
            nv1 = mesh.n[3];
  const long long nx = n_x_cells;
  const cptr<double> vfac = (cptr<double>)(tab_in + TB::vfac);
  const long long n_vfac = (long long)(nv0 + nv1) * 2 * N2;
  const cptr<double> vflag = vfac + n_vfac;
  const cptr<float> vfacf = (cptr<float>)(vflag + nv0 + nv1);

  auto vcoef = [&](long long i) -> T {
    if constexpr (sizeof(T) == 4)
      return vfacf[i];
    else
      return vfac[i];
  };

  __shared__ __align__(16) T lds[CPW * SLOT];
  // the lane and what derives from it are made opaque once per group, so
  // that the compiler recomputes the lane-dependent addresses in every
  // group instead of holding dozens of them in registers across the loop
  int lane;
  bool busy;
  int s, t;
  T *B0, *B1, *FH;
  double *D;
  constexpr int DSLOT = SLOT * (int)sizeof(T) / 8; // the density view
  auto lane_setup = [&]() {
    lane = (int)threadIdx.x;
    asm volatile("" : "+v"(lane));
    busy = G::LANES == 64 || lane < G::LANES; // idle lanes only load
    s = busy ? lane / N2 : 0;
    t = lane % N2;
    B0 = lds + s * SLOT;
    B1 = B0 + N2 * P;
    FH = B1; // the v faces at the x Gauss points, X1b -> V2
    // doubles over B0 and B1 of each slot (the density, FP64)
    D = reinterpret_cast<double *>(lds) + s * DSLOT;
  };
  lane_setup();

  // XCD-chunked persistent traversal: blocks are dispatched round-robin
  // over the 8 XCDs; each XCD walks a contiguous range of groups
  const long long bpx = gridDim.x / 8;
  const long long per = (n_groups + 7) / 8;
  const long long xcd = blockIdx.x % 8;
  const long long g_end = std::min(n_groups, (xcd + 1) * per);

  // coalesced value k of a lane (group element e = lane + 64 k: slot, then
  // x node fastest within a v node): its LDS offset, a lane part plus a
  // compile-time part when N^2 divides 64
  // (slot size sl: SLOT elements of T, or DSLOT doubles for the density)
  auto lofs = [&](int k, int sl) -> int {
    if constexpr (64 % N2 == 0 && N4 % 64 == 0)
      {
        constexpr int M = N4 / 64;
        return (k / M) * sl + (64 / N2) * (k % M) * P + (lane / N2) * P +
               lane % N2;
      }
    else if constexpr (64 % N4 == 0)
      return (lane / N4 + (64 / N4) * k) * sl + (lane % N4) / N2 * P +
             lane % N2;
    else
      {
        const int e = lane + 64 * k, ss = e / N4, tt = e % N4;
        return ss * sl + (tt / N2) * P + tt % N2;
      }
  };
  // f(k, e) for the lane's values of a group with n_in cells (a full group
  // without per-value checks)
  auto for_values = [&](int n_in, auto f) {
    if (n_in == CPW)
      {
#pragma unroll
        for (int k = 0; k < KU; ++k)
          {
            const int e = lane + 64 * k;
            if (64 * KU == CPW * N4 || e < CPW * N4)
              f(k, e);
          }
      }
    else
      {
#pragma unroll
        for (int k = 0; k < KU; ++k)
          {
            const int e = lane + 64 * k;
            if (e < n_in * N4)
              f(k, e);
          }
      }
  };

  // the next group's values, field at this lane's x Gauss point and x
  // inflow faces (the sign of pass 0) travel into registers early
  T pu[KU], pe[2], pf[2][N];
  auto prefetch = [&](long long g) {
    const long long vc = g / x_groups;
    const int xs = (int)(g - vc * x_groups) * CPW;
    const int n_in = (int)std::min<long long>(CPW, nx - xs);
    const long long base = (vc * nx + xs) * N4;
    for_values(n_in, [&](int k, int e) { pu[k] = src[base + e]; });
    const bool act = busy && s < n_in;
    const int xc = xs + s;
    const long long c = vc * nx + xc;
#pragma unroll
    for (int e = 0; e < 2; ++e)
      pe[e] = act ? (T)efield[((long long)xc * N2 + t) * 2 + e] : (T)0;
    const int cv0 = (int)(vc % nv0), cv1 = (int)(vc / nv0);
    const int f[2] = {(int)vflag[cv0], (int)vflag[nv0 + cv1]};
    const int cx0 = xc % n0, cx1 = xc / n0;
#pragma unroll
    for (int d = 0; d < 2; ++d)
      {
        const int sg = (f[d] & 1) ? 0 : 1;
        const long long nb = d == 0 ? nbr(c, cx0, n0, 1, sg)
                                    : nbr(c, cx1, n1, n0, sg);
        const int node = sg == 0 ? N - 1 : 0;
#pragma unroll
        for (int m = 0; m < N; ++m)
          {
            const int idx = d == 0 ? node + N * m : m + N * node;
            pf[d][m] = act && f[d] ? src[nb * N4 + idx + N2 * t] : (T)0;
          }
      }
  };

  long long g = xcd * per + blockIdx.x / 8;
  if (g < g_end)
    prefetch(g);
  for (; g < g_end; g += bpx)
    {
      tab = launder(tab0);
      tabf = (cptr<float>)(tab + TB::size);
      lane_setup();
      const long long vc = g / x_groups;
      const int xs = (int)(g - vc * x_groups) * CPW;
      const int n_in = (int)std::min<long long>(CPW, nx - xs);
      const bool act = busy && s < n_in;
      const int xc = xs + s;
      const long long c = vc * nx + xc;
      const long long base = (vc * nx + xs) * N4;
      const int cv0 = (int)(vc % nv0), cv1 = (int)(vc / nv0);
      const int f0 = (int)vflag[cv0], f1 = (int)vflag[nv0 + cv1];
      T ee[2] = {pe[0], pe[1]};
      T fx[2][N];
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int m = 0; m < N; ++m)
          fx[d][m] = pf[d][m];

      // the v faces this lane's cell needs: side 0 (left neighbour) where
      // some E_e >= 0, side 1 where some E_e < 0
      bool need[2][2];
      {
        const unsigned long long sm =
          busy ? (((1ull << N2) - 1) << (s * N2)) : 0ull;
#pragma unroll
        for (int e = 0; e < 2; ++e)
          {
            const unsigned long long ge = __ballot(act && ee[e] >= (T)0);
            const unsigned long long lt = __ballot(act && ee[e] < (T)0);
            need[e][0] = (ge & sm) != 0;
            need[e][1] = (lt & sm) != 0;
          }
      }
      T pv[RP][N2];
      bool have[RP];
#pragma unroll
      for (int r = 0; r < RP; ++r)
        {
          const int p = t + N2 * r;
          const int e = p / (2 * N), side = (p / N) % 2, m = p % N;
          have[r] = act && p < NP &&
                    (e == 0 ? (side == 0 ? need[0][0] : need[0][1])
                            : (side == 0 ? need[1][0] : need[1][1]));
          long long nb = 0;
          int off = 0;
          if (have[r])
            {
              const int node = side == 0 ? N - 1 : 0;
              nb = e == 0 ? nbr(c, cv0, nv0, nx, side)
                          : nbr(c, cv1, nv1, nx * nv0, side);
              off = e == 0 ? (node + N * m) * N2 : (m + N * node) * N2;
            }
#pragma unroll
          for (int i = 0; i < N2; ++i)
            pv[r][i] = have[r] ? src[nb * N4 + off + i] : (T)0;
        }
      // stage: the cells into B1 (x node fastest in a row of stride P per
      // v node), then this lane's x-plane into registers
      wave_sync();
      for_values(n_in, [&](int k, int) { lds[lofs(k, SLOT) + N2 * P] = pu[k]; });
      wave_sync();
      T u[N2];
#pragma unroll
      for (int i = 0; i < N2; ++i)
        u[i] = B1[t * P + i];

      // x terms, pass 0 (each direction's first sign) and, in a v cell
      // that contains v_d = 0, pass 1 (the minus sign)
      T nodal[N2];
#pragma unroll
      for (int j = 0; j < N2; ++j)
        nodal[j] = 0;
#pragma unroll
      for (int pass = 0; pass < 2; ++pass)
        {
          const int sg0 = pass == 0 ? ((f0 & 1) ? 0 : (f0 & 2) ? 1 : -1)
                                    : (f0 == 3 ? 1 : -1);
          const int sg1 = pass == 0 ? ((f1 & 1) ? 0 : (f1 & 2) ? 1 : -1)
                                    : (f1 == 3 ? 1 : -1);
          if (pass == 1 && sg0 < 0 && sg1 < 0)
            break;
          if (pass == 1)
            {
              // the right faces (node 0 of the right neighbours)
              const int cx0 = xc % n0, cx1 = xc / n0;
#pragma unroll
              for (int d = 0; d < 2; ++d)
                {
                  const bool on = d == 0 ? sg0 >= 0 : sg1 >= 0;
                  const long long nb = d == 0 ? nbr(c, cx0, n0, 1, 1)
                                              : nbr(c, cx1, n1, n0, 1);
#pragma unroll
                  for (int m = 0; m < N; ++m)
                    {
                      const int idx = d == 0 ? N * m : m;
                      fx[d][m] =
                        act && on ? src[nb * N4 + idx + N2 * t] : (T)0;
                    }
                }
            }
          // X1: P_d = A_d u + b_d face along x_d, sign sg_d
          if (busy && sg0 >= 0)
            {
              const int oA = sg0 == 0 ? TB::Ap(0) : TB::An(0);
              const int ob = sg0 == 0 ? TB::bp(0) : TB::bn(0);
              // difference form (A^ 1 = 0): A^ u + b (f - u_o); FP32 applies
              // A^ to u - u_o as well, so that no product carries the
              // magnitude of u
              T dd[N2], uo[N];
#pragma unroll
              for (int b = 0; b < N; ++b)
                {
                  uo[b] = sg0 == 0 ? u[N - 1 + N * b] : u[N * b];
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    dd[k + N * b] =
                      sizeof(T) == 4 ? u[k + N * b] - uo[b] : u[k + N * b];
                }
              T y[N2];
              along0<N>(dd, y, [&](int r) { return tc(oA + r); });
#pragma unroll
              for (int b = 0; b < N; ++b)
                {
                  const T df = fx[0][b] - uo[b];
#pragma unroll
                  for (int i = 0; i < N; ++i)
                    {
                      T v = y[i + N * b] + tc(ob + i) * df;
                      B0[t * P + i + N * b] = v;
                    }
                }
            }
          if (busy && sg1 >= 0)
            {
              const int oA = sg1 == 0 ? TB::Ap(1) : TB::An(1);
              const int ob = sg1 == 0 ? TB::bp(1) : TB::bn(1);
              T dd[N2], uo[N];
#pragma unroll
              for (int a = 0; a < N; ++a)
                {
                  uo[a] = sg1 == 0 ? u[a + N * (N - 1)] : u[a];
#pragma unroll
                  for (int k = 0; k < N; ++k)
                    dd[a + N * k] =
                      sizeof(T) == 4 ? u[a + N * k] - uo[a] : u[a + N * k];
                }
              T y[N2];
              along1<N>(dd, y, [&](int r) { return tc(oA + r); });
#pragma unroll
              for (int a = 0; a < N; ++a)
                {
                  const T df = fx[1][a] - uo[a];
#pragma unroll
                  for (int i = 0; i < N; ++i)
                    {
                      T v = y[a + N * i] + tc(ob + i) * df;
                      B1[t * P + a + N * i] = v;
                    }
                }
            }
          wave_sync();
          // V1: nodal += V_0 P_0 (along v0) + V_1 P_1 (along v1), lane =
          // x node t
          if (busy && sg0 >= 0)
            {
              const long long o = ((long long)cv0 * 2 + sg0) * N2;
              T q[N2];
#pragma unroll
              for (int j = 0; j < N2; ++j)
                q[j] = B0[j * P + t];
              along0_acc<N>(q, nodal, [&](int r) { return vcoef(o + r); });
            }
          if (busy && sg1 >= 0)
            {
              const long long o =
                (long long)nv0 * 2 * N2 + ((long long)cv1 * 2 + sg1) * N2;
              T q[N2];
#pragma unroll
              for (int j = 0; j < N2; ++j)
                q[j] = B1[j * P + t];
              along1_acc<N>(q, nodal, [&](int r) { return vcoef(o + r); });
            }
          wave_sync();
        }

      // X1b: the cell and the v faces to the x Gauss points
      if (busy)
        {
          T y[N2], z[N2];
          along0<N>(u, y, [&](int r) { return tc(TB::S + r); });
          along1<N>(y, z, [&](int r) { return tc(TB::S + r); });
#pragma unroll
          for (int q = 0; q < N2; ++q)
            B0[t * P + q] = z[q];
#pragma unroll
          for (int r = 0; r < RP; ++r)
            if (have[r])
              {
                const int p = t + N2 * r;
                along0<N>(pv[r], y, [&](int rr) { return tc(TB::S + rr); });
                along1<N>(y, z, [&](int rr) { return tc(TB::S + rr); });
#pragma unroll
                for (int q = 0; q < N2; ++q)
                  FH[p * P + q] = z[q];
              }
        }
      wave_sync();

      // V2: the v operators at this lane's x Gauss point t; FH plane
      // (e, side, m) = (e * 2 + side) * N + m
      if (busy)
        {
          T uh[N2], w[N2];
#pragma unroll
          for (int j = 0; j < N2; ++j)
            {
              uh[j] = B0[j * P + t];
              w[j] = 0;
            }
          // w += c (A u^ + b face^) along v_e for one upwind side; masked:
          // the face only where this lane's E_e selects that side
          auto part = [&](auto e_c, auto side_c, T c, bool masked) {
            constexpr int e = decltype(e_c)::value;
            constexpr int side = decltype(side_c)::value;
            constexpr int oA = side == 0 ? TB::Ap(2 + e) : TB::An(2 + e);
            constexpr int ob = side == 0 ? TB::bp(2 + e) : TB::bn(2 + e);
            const bool use = side == 0 ? ee[e] >= (T)0 : ee[e] < (T)0;
            const T *fh = FH + (2 * e + side) * N * P + t;
            // difference form along v_e (FP32: lines minus their outflow
            // value)
            constexpr int o = side == 0 ? N - 1 : 0;
            T dd[N2], uo[N];
#pragma unroll
            for (int m = 0; m < N; ++m)
              {
                uo[m] = e == 0 ? uh[o + N * m] : uh[m + N * o];
#pragma unroll
                for (int k = 0; k < N; ++k)
                  {
                    const int j = e == 0 ? k + N * m : m + N * k;
                    dd[j] = sizeof(T) == 4 ? uh[j] - uo[m] : uh[j];
                  }
              }
            T y[N2];
            if constexpr (e == 0)
              along0<N>(dd, y, [&](int r) { return tc(oA + r); });
            else
              along1<N>(dd, y, [&](int r) { return tc(oA + r); });
#pragma unroll
            for (int m = 0; m < N; ++m)
              {
                T df = fh[m * P] - uo[m];
                if (masked && !use)
                  df = 0;
#pragma unroll
                for (int i = 0; i < N; ++i)
                  {
                    T v = tc(ob + i) * df;
                    y[e == 0 ? i + N * m : m + N * i] += v;
                  }
              }
#pragma unroll
            for (int j = 0; j < N2; ++j)
              w[j] += c * y[j];
          };
          // one branch for the wave where its lanes agree on the sign
          auto vdir = [&](auto e_c) {
            constexpr int e = decltype(e_c)::value;
            using I0 = std::integral_constant<int, 0>;
            using I1 = std::integral_constant<int, 1>;
            const T c = ee[e];
            const unsigned long long neg = __ballot(c < (T)0);
            const unsigned long long pos = __ballot(c >= (T)0);
            if (neg == 0)
              part(e_c, I0(), c, false);
            else if (pos == 0)
              part(e_c, I1(), c, false);
            else
              {
                part(e_c, I0(), c >= (T)0 ? c : (T)0, true);
                part(e_c, I1(), c < (T)0 ? c : (T)0, true);
              }
          };
          vdir(std::integral_constant<int, 0>());
          vdir(std::integral_constant<int, 1>());
#pragma unroll
          for (int j = 0; j < N2; ++j)
            {
              B0[j * P + t] = w[j];
              B1[j * P + t] = nodal[j];
            }
        }
      wave_sync();

      // X2: back to the nodes along x0, x1, plus the x terms
      if (busy)
        {
          T wq[N2], y[N2], z[N2];
#pragma unroll
          for (int q = 0; q < N2; ++q)
            wq[q] = B0[t * P + q];
          along0<N>(wq, y, [&](int r) { return tc(TB::Si + r); });
          along1<N>(y, z, [&](int r) { return tc(TB::Si + r); });
#pragma unroll
          for (int i = 0; i < N2; ++i)
            B0[t * P + i] = z[i] + B1[t * P + i];
        }
      wave_sync();
      // the solution (LSRK stage)
      double ps[KU];
      if constexpr (MODE == 1)
        for_values(n_in, [&](int k, int e) { ps[k] = stage.sol[base + e]; });

      // out: coalesced, x node fastest; FP64 keeps the unrounded stage value
      // in place for the density, FP32 in registers until all are read
      const bool dens = MODE == 1 && stage.rho_part && stage.ti_out;
      constexpr int KT = sizeof(T) == 4 ? KU : 1;
      double tnk[KT];
      for_values(n_in, [&](int k, int e) {
        const T a = lds[lofs(k, SLOT)];
        if constexpr (MODE == 0)
          dst[base + e] = a;
        else
          {
            const double so = ps[k];
            stage.sol[base + e] = so + stage.b_dt * (double)a;
            if (stage.ti_out)
              {
                const double tn = so + stage.a_dt * (double)a;
                static_cast<T *>(stage.ti_out)[base + e] = (T)tn;
                if constexpr (sizeof(T) == 4)
                  tnk[k] = tn;
                else if (dens)
                  lds[lofs(k, SLOT)] = tn;
              }
          }
      });
      // the next group's loads go out now and land during the density and
      // the next group's staging
      if (g + bpx < g_end)
        prefetch(g + bpx);
      if (dens)
        {
          // int f dv per x node (the Gauss-Lobatto v rule on the nodes,
          // numerics/vector_tools.h:230-315) in FP64 from the unrounded
          // stage value: one partial per (v cell, x node)
          if constexpr (sizeof(T) == 4)
            {
              wave_sync(); // the double view overlaps other lanes' values
              for_values(n_in, [&](int k, int) {
                reinterpret_cast<double *>(lds)[lofs(k, DSLOT)] = tnk[k];
              });
            }
          wave_sync();
          if (act)
            {
              double acc = 0.0;
#pragma unroll
              for (int j = 0; j < N2; ++j)
                acc += tab[TB::wv + j] * D[j * P + t];
              stage.rho_part[(vc * nx + xc) * N2 + t] = acc;
            }
        }
    }
}

// resident waves per CU of a kernel (one wave per workgroup), cached
template <int N, typename T, int MODE>
int waves_per_cu()
{
  static int w = 0;
  if (w == 0)
    {
      int b = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b, stage_kernel<N, T, MODE>, 64, 0) != hipSuccess ||
          b < 1)
        b = 1;
      w = b;
    }
  return w;
}

template <int N, typename T>
hipError_t launch(int mode, const T *src, T *dst, const double *tab,
                  const MeshDev &mesh, long long n_x_cells,
                  const double *efield, const StageDev &stage,
                  hipStream_t stream)
{
  using G = Geo<N>;
  const long long x_groups = (n_x_cells + G::CPW - 1) / G::CPW;
  const long long n_groups = x_groups * (mesh.n_cells / n_x_cells);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess ||
      cus < 8)
    cus = 256;
  const int wpc =
    mode == 0 ? waves_per_cu<N, T, 0>() : waves_per_cu<N, T, 1>();
  // a multiple of 8 blocks (the XCD count), at most one round of groups
  long long grid = (long long)cus * wpc;
  grid = std::min<long long>(grid, (n_groups + 7) / 8 * 8);
  grid = std::max<long long>(8, grid / 8 * 8);
  if (mode == 0)
    hipLaunchKernelGGL((stage_kernel<N, T, 0>), dim3((unsigned)grid),
                       dim3(64), 0, stream, src, dst, tab, mesh, n_x_cells,
                       efield, stage, x_groups, n_groups);
  else
    hipLaunchKernelGGL((stage_kernel<N, T, 1>), dim3((unsigned)grid),
                       dim3(64), 0, stream, src, dst, tab, mesh, n_x_cells,
                       efield, stage, x_groups, n_groups);
  return hipGetLastError();
}
} // namespace vpw
} // namespace vlasov

hipError_t vp_wave_apply(int N, int mode, int precision, const void *src,
                         void *dst, const double *tab, const MeshDev &mesh,
                         long long n_x_cells, const double *efield,
                         const StageDev &stage, hipStream_t stream)
{
#define HD_VPW(NN_)                                                           \
  if (N == NN_)                                                               \
    return precision == 0                                                     \
             ? vlasov::vpw::launch<NN_, double>(                              \
                 mode, (const double *)src, (double *)dst, tab, mesh,         \
                 n_x_cells, efield, stage, stream)                            \
             : vlasov::vpw::launch<NN_, float>(                               \
                 mode, (const float *)src, (float *)dst, tab, mesh,           \
                 n_x_cells, efield, stage, stream);
  HD_VPW(2) HD_VPW(3) HD_VPW(4) HD_VPW(5)
#undef HD_VPW
  return hipErrorInvalidValue;
}
} // namespace hd
