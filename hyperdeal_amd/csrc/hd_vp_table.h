// hd_vp_table.h -- the coefficient table of the 2+2 Vlasov-Poisson apply
// (shared by the phase-space kernel and its host-side builder).
//
// Doubles, in order:
//   per direction d (x0, x1, v0, v1): A+ (N^2), b+ (N), A- (N^2), b- (N) --
//     the collapsed 1D operator for speed +1 and minus the one for speed -1,
//     so that the operator for speed c is c A+ (c >= 0) or c A- (c < 0),
//     with A stored in difference form A^ = A + b e_o^T (o: the outflow
//     node, N-1 for +, 0 for -): A u + b f = A^ (u - u_o) + b (f - u_o);
//   S (N^2, row q: Gauss point), S^-1 (N^2), the Gauss points (N), JxW_v at
//     the Gauss-Lobatto v nodes (N^2);
// then the FP32 copy of all of the above (size / 2 rounded up doubles);
// then the v factors of the x terms S^-1 diag(v_d(q)+/-) S per v_d cell,
//   [d][v_d cell][sign][i][l] (N^2 each);
// then per d and v_d cell a flag: bit 0 set when the + factor is nonzero
//   (some Gauss point has v_d > 0), bit 1 when the - factor is;
// then the FP32 copy of the v factors.
#pragma once

namespace hd
{
namespace vlasov
{
namespace vptab
{
template <int N>
struct Tab
{
  static constexpr int NN = N * N, PER = 2 * (NN + N);
  static constexpr int Ap(int d) { return d * PER; }
  static constexpr int bp(int d) { return d * PER + NN; }
  static constexpr int An(int d) { return d * PER + NN + N; }
  static constexpr int bn(int d) { return d * PER + 2 * NN + N; }
  static constexpr int S = 4 * PER, Si = S + NN, qp = Si + NN;
  static constexpr int wv = qp + N; // JxW_v of the v nodes (N^2)
  static constexpr int size = wv + NN;
  // offset of the v factors (doubles) behind the FP32 copy
  static constexpr int vfac = size + (size + 1) / 2;
};
} // namespace vptab
} // namespace vlasov
} // namespace hd
