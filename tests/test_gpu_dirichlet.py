"""Dirichlet boundaries through the collapsed kernels (the periodic-image
kernel followed by the boundary correction of hd_dirichlet.hip) against the
CPU oracle, which restates the reference's ECL quadrature algorithm with the
exterior value u+ = -u- + 2 g at boundary faces (advection_operation.h:
484-520; u+ = -u- for homogeneous data).

Bars (per test):
  - FP64 apply: max |gpu - oracle| <= 1e-12 max |oracle|;
  - FP32 apply: max |gpu - oracle(fp32-rounded input)| <= 2e-6 max |oracle|;
  - fused LSRK stages (middle and last) equal sol += b dt K and
    ti = sol_old + a dt K with K from the collapsed apply, to 1e-13 (FP64)
    and 2e-6 of b dt max |K| (FP32-mixed).
The cases cover every collapsed kernel family: the wave-per-group 2+2
kernel (k = 2, 3; FP32 k = 4), the general kernel (2+2 k = 4 FP64, k = 5;
1+1; 2+3), the 3+3 k=3 row walk, the 3+3 k = 4, 5 split form (FP64) and
one-pass kernel (FP32), Dirichlet in all or some directions, over-integrated
face quadrature, a direction without transport and homogeneous data.
"""

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

H = None


def setup_module(module):
    global H
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hyperdeal_amd
    H = hyperdeal_amd
    H.load_library()


CASES = [
    # dims, degree, n_cells, periodic_x, periodic_v, n_points, bc, zero dir
    ((1, 1), 3, [4, 3], False, False, None, O.BC_INHOMOGENEOUS, None),
    ((2, 2), 2, [3, 2, 2, 3], False, False, None, O.BC_INHOMOGENEOUS, None),
    ((2, 2), 3, [3, 2, 2, 3], True, False, None, O.BC_INHOMOGENEOUS, None),
    ((2, 2), 3, [2, 3, 2, 2], False, False, None, O.BC_HOMOGENEOUS, 1),
    ((2, 2), 4, [2, 3, 2, 2], False, False, None, O.BC_INHOMOGENEOUS, None),
    ((2, 2), 5, [2, 2, 2, 2], False, True, None, O.BC_INHOMOGENEOUS, None),
    ((2, 2), 2, [3, 2, 2, 2], False, False, 5, O.BC_INHOMOGENEOUS, None),
    ((3, 3), 3, [2, 3, 2, 2, 2, 2], False, False, None, O.BC_INHOMOGENEOUS,
     None),
    ((3, 3), 3, [2, 2, 2, 2, 2, 2], False, False, 5, O.BC_INHOMOGENEOUS, 4),
    ((3, 3), 5, [2, 1, 1, 2, 1, 1], False, False, None, O.BC_INHOMOGENEOUS,
     None),
    ((3, 3), 4, [1, 2, 1, 1, 2, 1], True, False, None, O.BC_INHOMOGENEOUS,
     None),
    ((2, 3), 4, [2, 2, 1, 2, 1], False, False, None, O.BC_INHOMOGENEOUS,
     None),
]
IDS = ["%d+%d_k%d_%s%s%s%s" % (c[0][0], c[0][1], c[1],
                               "P" if c[3] else "D", "P" if c[4] else "D",
                               "_q%d" % c[5] if c[5] else "",
                               "_hom" if c[6] == O.BC_HOMOGENEOUS else "")
       for c in CASES]


def _setup(case, seed):
    dims, k, n_cells, px, pv, nq, bc, zero = case
    D = sum(dims)
    r = np.random.default_rng(seed)
    a = list(r.uniform(-1, 1, D))
    if zero is not None:
        a[zero] = 0.0
    lo = list(r.uniform(-0.5, 0.0, D))
    hi = [l + w for l, w in zip(lo, r.uniform(0.5, 2.0, D))]
    P = O.Problem(dims[0], dims[1], k, n_cells, lo, hi, a, n_points=nq,
                  periodic_x=px, periodic_v=pv, bc=bc,
                  func=O.FUNC_HYPERRECTANGLE)
    mf = H.MatrixFree(dims[0], dims[1], k, n_cells, lo, hi, periodic_x=px,
                      periodic_v=pv, n_points=nq)
    op = H.AdvectionOperation(mf, a, boundary=bc,
                              boundary_function=H.FUNC_HYPERRECTANGLE,
                              boundary_params=a)
    return P, mf, op


def _rel(x, ref):
    return np.max(np.abs(x - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dirichlet_collapsed_apply_matches_oracle(case, precision):
    print("case: dirichlet collapsed apply", case[:3], precision, flush=True)
    P, mf, op = _setup(case, seed=3)
    assert op.algorithm() == H.ALGO_COLLAPSED
    u = np.random.default_rng(4).standard_normal(P.n_dofs)
    for t in (0.0, 0.37):
        if precision == "fp64":
            ref = P.apply(u, t=t)
            src = torch.from_numpy(u).cuda()
            dst = torch.empty_like(src)
            op.apply(dst, src, time=t)
            torch.cuda.synchronize()
            assert _rel(dst.cpu().numpy(), ref) <= 1e-12, t
        else:
            u32 = u.astype(np.float32)
            ref = P.apply(u32.astype(np.float64), t=t)
            src = torch.from_numpy(u32).cuda()
            dst = torch.empty_like(src)
            op.apply(dst, src, time=t, precision=H.FP32)
            torch.cuda.synchronize()
            assert _rel(dst.cpu().numpy().astype(np.float64), ref) <= 2e-6, t


@pytest.mark.parametrize("case", [CASES[1], CASES[7], CASES[9], CASES[11]],
                         ids=[IDS[1], IDS[7], IDS[9], IDS[11]])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dirichlet_collapsed_stages(case, precision):
    """The fused stages with the boundary correction added to sol and ti."""
    print("case: dirichlet collapsed stages", case[:3], precision, flush=True)
    P, mf, op = _setup(case, seed=5)
    r = np.random.default_rng(6)
    dt = torch.float64 if precision == "fp64" else torch.float32
    prec = H.FP64 if precision == "fp64" else H.FP32
    u = torch.from_numpy(r.standard_normal(P.n_dofs)).cuda().to(dt)
    sol0 = torch.from_numpy(r.standard_normal(P.n_dofs)).cuda()
    t, a_dt, b_dt = 0.21, 0.013, 0.029
    k = torch.empty_like(u)
    op.apply(k, u, time=t, precision=prec)
    torch.cuda.synchronize()
    k64 = k.double()
    kmax = k64.abs().max().item()
    # middle stage
    sol = sol0.clone()
    ti = torch.empty_like(u)
    op.stage(u, ti, sol, t, a_dt, b_dt, precision=prec)
    torch.cuda.synchronize()
    tol = 1e-13 if precision == "fp64" else 2e-6
    assert (sol - (sol0 + b_dt * k64)).abs().max().item() <= \
        tol * b_dt * kmax + 1e-15 * sol0.abs().max().item()
    ti_ref = sol0 + a_dt * k64
    assert (ti.double() - ti_ref).abs().max().item() <= \
        tol * a_dt * kmax + (6e-8 if precision == "fp32" else 1e-15) * \
        ti_ref.abs().max().item()
    # last stage (no ti_out)
    sol = sol0.clone()
    op.stage(u, None, sol, t, 0.0, b_dt, precision=prec)
    torch.cuda.synchronize()
    assert (sol - (sol0 + b_dt * k64)).abs().max().item() <= \
        tol * b_dt * kmax + 1e-15 * sol0.abs().max().item()


def test_dirichlet_collapsed_equals_quadrature_kernel():
    """Both GPU forms of a Dirichlet operator on a larger mesh (2+2 k=3,
    6^4 cells): the collapsed kernel with the boundary correction and the
    quadrature kernel (the reference's steps) agree to 1e-12."""
    print("case: dirichlet collapsed vs quadrature", flush=True)
    D = 4
    a = [0.6, -0.45, 0.3, -0.8]
    mf = H.MatrixFree(2, 2, 3, [6] * D, [-1.0] * D, [1.0] * D,
                      periodic_x=False, periodic_v=False)
    op = H.AdvectionOperation(mf, a, boundary=H.BC_INHOMOGENEOUS,
                              boundary_function=H.FUNC_HYPERRECTANGLE,
                              boundary_params=a)
    u = torch.from_numpy(
        np.random.default_rng(8).standard_normal(mf.n_dofs())).cuda()
    out = {}
    for algo in (H.ALGO_COLLAPSED, H.ALGO_QUADRATURE):
        op.set_algorithm(algo)
        dst = torch.empty_like(u)
        op.apply(dst, u, time=0.29)
        out[algo] = dst
    torch.cuda.synchronize()
    ref = out[H.ALGO_QUADRATURE].cpu().numpy()
    assert _rel(out[H.ALGO_COLLAPSED].cpu().numpy(), ref) <= 1e-12


def test_dirichlet_batches_rejected():
    """Batches run periodic operators only: a Dirichlet operator is an
    error when the batch is created."""
    mf = H.MatrixFree(2, 2, 2, [2] * 4, [0.0] * 4, [1.0] * 4,
                      periodic_x=False, periodic_v=False)
    op = H.AdvectionOperation(mf, [0.3, 0.2, -0.1, 0.4])
    with pytest.raises(H.HyperdealError):
        H.Batch([op])
