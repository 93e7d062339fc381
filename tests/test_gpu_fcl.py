"""The face-centric (FCL) form on the GPU: MatrixFree::loop's cell, face and
boundary operations and the inverse-mass cell_loop
(matrix_free.templates.h:1417-1465, 1634-1746; advection_operation.h:
159-183, 573-1131), through the C ABI with HD_ALGO_FCL, against the
oracle's FCL.

Bars (written in each test):
  - FP64 apply: max |gpu - oracle FCL| <= 1e-12 max |oracle|, and the same
    against the oracle's ECL (the two forms are the same operator)
  - FP32 apply: 2e-6 of max |oracle| on the FP32-rounded input
  - LSRK steps (all stages, the last one assembling in the operator's own
    vector): 1e-12 against the oracle's FCL steps
  - the reference's FCL goldens (Matrixfree/UseECL = false) end to end on
    the GPU through the FCL kernels: numdiff -a 1e-5 -r 1e-8
Meshes with an odd number of cells in a periodic direction use the third
face colour (the wrap face); one cell in a periodic direction puts both
sides of a face in the same cell.
"""

import numpy as np
import pytest

import oracle as O
from tests.golden import advection_cases as AC

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


def _pair(dims, degree, n_cells, seed, n_points=None, collocation=False,
          periodic=True, skew=0.0, bc=O.BC_INHOMOGENEOUS):
    D = sum(dims)
    r = np.random.default_rng(seed)
    a = list(r.uniform(-1, 1, D))
    lo = [-1.0] * D
    hi = list(r.uniform(0.5, 2.0, D))
    func = O.FUNC_HYPERRECTANGLE
    P = O.Problem(dims[0], dims[1], degree, n_cells, lo, hi, a,
                  n_points=n_points, collocation=collocation,
                  periodic_x=periodic, periodic_v=periodic, skew=skew, bc=bc,
                  func=func)
    mf = H.MatrixFree(dims[0], dims[1], degree, n_cells, lo, hi,
                      periodic_x=periodic, periodic_v=periodic,
                      n_points=n_points, collocation=collocation)
    op = H.AdvectionOperation(mf, a, skew_factor=skew, boundary=bc,
                              boundary_function=func, boundary_params=a)
    op.set_algorithm(H.ALGO_FCL)
    assert op.algorithm() == H.ALGO_FCL
    return P, mf, op, a


def _rel_err(x, ref):
    return np.max(np.abs(x - ref)) / np.max(np.abs(ref))


CASES = [
    # dims, degree, n_cells, periodic, skew, n_points, collocation
    ((1, 1), 3, [5, 4], True, 0.0, None, False),
    ((1, 1), 1, [1, 3], True, 0.0, None, False),
    ((1, 1), 3, [4, 3], False, 0.5, None, False),
    ((2, 2), 3, [3, 2, 3, 2], True, 0.0, None, False),
    ((2, 2), 3, [3, 2, 2, 3], False, 0.0, None, False),
    ((2, 2), 2, [3, 3, 2, 2], False, 1.0, None, False),
    ((2, 2), 5, [2, 2, 2, 3], True, 0.5, None, False),
    ((2, 2), 3, [3, 2, 2, 3], False, 0.0, None, True),
    ((2, 2), 3, [2, 3, 1, 2], True, 0.0, None, True),
    ((2, 2), 3, [3, 2, 2, 3], False, 0.0, 5, False),
    ((2, 2), 4, [2, 2, 2, 2], True, 0.3, 6, False),
    ((1, 2), 3, [3, 2, 2], False, 0.0, None, False),
    ((2, 1), 4, [2, 3, 2], True, 0.0, None, False),
    ((3, 2), 3, [2, 2, 2, 2, 3], True, 0.0, None, False),
    ((3, 3), 3, [2, 3, 2, 2, 3, 2], True, 0.0, None, False),
    ((3, 3), 2, [3, 2, 2, 2, 2, 3], False, 0.5, None, False),
    ((3, 3), 2, [2, 2, 2, 2, 2, 2], False, 0.0, 4, False),
]


def _ids(cases):
    return ["%d+%d_k%d_%s_s%g%s%s" % (d[0], d[1], k, "per" if p else "dir",
                                      s, "_q%d" % q if q else "",
                                      "_coll" if c else "")
            for d, k, _, p, s, q, c in cases]


@pytest.mark.parametrize("dims,degree,n_cells,periodic,skew,n_points,coll",
                         CASES, ids=_ids(CASES))
def test_fcl_apply_matches_oracle(dims, degree, n_cells, periodic, skew,
                                  n_points, coll):
    print("case: fcl apply", dims, degree, n_cells, periodic, skew, n_points,
          coll, flush=True)
    P, mf, op, a = _pair(dims, degree, n_cells, seed=degree * 10 + sum(dims),
                         n_points=n_points, collocation=coll,
                         periodic=periodic, skew=skew)
    u = np.random.default_rng(1).standard_normal(P.n_dofs)
    src = torch.from_numpy(u).cuda()
    for t in (0.0, 0.37):
        ref_fcl = P.apply(u, t=t, fcl=True)
        ref_ecl = P.apply(u, t=t, fcl=False)
        dst = torch.empty_like(src)
        op.apply(dst, src, time=t)
        torch.cuda.synchronize()
        got = dst.cpu().numpy()
        assert _rel_err(got, ref_fcl) <= 1e-12
        assert _rel_err(got, ref_ecl) <= 1e-12
    # FP32: float vectors and contractions
    u32 = src.float()
    ref = P.apply(u32.double().cpu().numpy(), fcl=True)
    d32 = torch.empty_like(u32)
    op.apply(d32, u32, precision=H.FP32)
    torch.cuda.synchronize()
    assert _rel_err(d32.double().cpu().numpy(), ref) <= 2e-6


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("rk", ["rk33", "rk45"])
def test_fcl_lsrk_steps_match_oracle(rk, periodic):
    print("case: fcl lsrk", rk, periodic, flush=True)
    P, mf, op, a = _pair((2, 2), 3, [3, 2, 2, 3], seed=14, periodic=periodic)
    u = np.random.default_rng(6).standard_normal(P.n_dofs)
    sol_ref, ki, ti = u.copy(), P.vector(), P.vector()
    dt = 0.01
    for step in range(3):
        P.lsrk_step(rk, sol_ref, ki, ti, step * dt, dt, ecl=False)
    sol = torch.from_numpy(u.copy()).cuda()
    ta, tb = torch.empty_like(sol), torch.empty_like(sol)
    integ = H.LowStorageRungeKuttaIntegrator(rk)
    for step in range(3):
        integ.perform_time_step(op, sol, ta, tb, step * dt, dt)
    torch.cuda.synchronize()
    assert _rel_err(sol.cpu().numpy(), sol_ref) <= 1e-12
    if not periodic:
        return
    # single stages, the last one without a next-stage register: the
    # operator assembles in a vector of its own (periodic: the operator
    # does not depend on the stage time)
    A, B = integ.ai, integ.bi
    s1 = torch.from_numpy(u.copy()).cuda()
    tin = s1.clone()
    tout = torch.empty_like(s1)
    ns = len(B)
    for s in range(ns):
        last = s == ns - 1
        op.stage(tin, None if last else tout, s1, 0.0,
                 0.0 if last else A[s] * dt, B[s] * dt)
        tin, tout = tout, tin
    s2 = torch.from_numpy(u.copy()).cuda()
    integ.perform_time_step(op, s2, ta, tb, 0.0, dt)
    torch.cuda.synchronize()
    assert _rel_err(s1.cpu().numpy(), s2.cpu().numpy()) <= 1e-13


GOLD_FCL = [c for c in AC.cases() if not c["use_ecl"]]


@pytest.mark.parametrize("c", GOLD_FCL, ids=[c["name"] for c in GOLD_FCL])
def test_reference_fcl_golden_on_gpu(c):
    """The reference's FCL runs (examples/advection/tests with
    Matrixfree/UseECL = false) end to end through the GPU FCL kernels."""
    print("case: fcl golden %s" % c["name"], flush=True)
    D = c["dim_x"] + c["dim_v"]
    lo, hi = AC.HYPERRECTANGLE_BOX
    a = AC.HYPERRECTANGLE_VELOCITY[:D]
    mf = H.MatrixFree(c["dim_x"], c["dim_v"], c["degree"], c["n_cells"],
                      [lo] * D, [hi] * D, periodic_x=c["periodic_x"],
                      periodic_v=c["periodic_v"], n_points=c["n_points"],
                      collocation=c["collocation"])
    op = H.AdvectionOperation(mf, a, skew_factor=c["skew"],
                              boundary=H.BC_INHOMOGENEOUS,
                              boundary_function=H.FUNC_HYPERRECTANGLE,
                              boundary_params=a)
    op.set_algorithm(H.ALGO_FCL)
    sol = mf.initialize_dof_vector()
    ta, tb = torch.empty_like(sol), torch.empty_like(sol)
    H.VectorTools.interpolate(mf, sol, H.FUNC_HYPERRECTANGLE, a,
                              c["start_time"])
    integ = H.LowStorageRungeKuttaIntegrator(c["rk"])
    crit = H.compute_critical_time_step(mf, a)
    dt = min(c["time_step"], c["cfl"] * crit / c["degree"] ** 1.5)
    t0, T, tick = c["start_time"], c["final_time"], c["tick"]
    dt = (T - t0) / np.ceil((T - t0) / dt)
    rows = []

    def diag(t):
        if t != t0 and int((t + 1e-11 - t0) / tick) == \
                int((t + 1e-11 - t0 - dt) / tick):
            return
        n, e = H.VectorTools.norm_and_error(mf, sol, H.FUNC_HYPERRECTANGLE,
                                            a, t)
        rows.append((t, n, e))

    diag(t0)
    t = t0 + dt
    while t <= T * 1.0000000000001:
        integ.perform_time_step(op, sol, ta, tb, t - dt, dt)
        diag(t)
        t += dt
    ours = AC.parse_rows(AC.format_rows(np.array(rows)))
    assert ours.shape == c["golden"].shape
    assert AC.numdiff_ok(ours, c["golden"])


def test_fcl_rejects_vlasov_poisson():
    mf = H.MatrixFree(2, 2, 3, [4, 4, 4, 4], [0, 0, -6, -6],
                      [4 * np.pi, 4 * np.pi, 6, 6])
    op = H.VlasovPoissonOperation(mf)
    with pytest.raises(H.HyperdealError):
        op.set_algorithm(H.ALGO_FCL)
