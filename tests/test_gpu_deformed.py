"""Deformed meshes on the GPU (the reference's DeformedCubeManifold,
source/grid/grid_generator.cc:110-230, as the exact map of each space):
per-point Jacobians, JxW and normals in the quadrature (ECL) and
face-centric (FCL) kernels, through the C ABI, against the oracle.

Bars (written in each test):
  - FP64 apply, ECL and FCL kernels: max |gpu - oracle| <= 1e-12 max |oracle|
    (the oracle evaluates J^-1, det J and the normals J^-T n^ / |J^-T n^|
    per point; the kernels use per-space tables of J^-1 a, |det J| w and
    (a . n) JxW_face)
  - FP32 apply: 2e-6 of max |oracle| on the FP32-rounded input
  - interpolation, L2 norm / error and the critical time step at the
    mapped points: 1e-12 relative
  - three LSRK steps against the oracle's: 1e-12
The oracle's deformed operator is pinned by invariants in
tests/test_oracle_deformed.py (no reference golden covers a deformed mesh).
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


def _pair(dims, k, n, periodic=True, skew=0.0, n_points=None,
          collocation=False, delta=0.1, freq=2, seed=0):
    D = sum(dims)
    r = np.random.default_rng(seed)
    a = list(r.uniform(-1, 1, D))
    lo, hi = [-1.0] * D, [1.0] * D
    P = O.Problem(dims[0], dims[1], k, n, lo, hi, a, n_points=n_points,
                  collocation=collocation, periodic_x=periodic,
                  periodic_v=periodic, skew=skew,
                  func=O.FUNC_HYPERRECTANGLE, func_params=a)
    P.set_deformation(delta, freq)
    mf = H.MatrixFree(dims[0], dims[1], k, n, lo, hi, periodic_x=periodic,
                      periodic_v=periodic, n_points=n_points,
                      collocation=collocation)
    mf.set_deformation(delta, freq)
    op = H.AdvectionOperation(mf, a, skew_factor=skew,
                              boundary=H.BC_INHOMOGENEOUS,
                              boundary_function=H.FUNC_HYPERRECTANGLE,
                              boundary_params=a)
    return P, mf, op, a


def _rel(x, ref):
    return np.max(np.abs(x - ref)) / np.max(np.abs(ref))


CASES = [
    # dims, degree, n, periodic, skew, n_points, collocation
    ((1, 1), 3, [5, 4], True, 0.0, None, False),
    ((1, 1), 2, [4, 3], False, 0.5, None, False),
    ((2, 2), 3, [3, 2, 2, 3], True, 0.0, None, False),
    ((2, 2), 3, [3, 2, 2, 3], False, 0.0, None, False),
    ((2, 2), 2, [3, 3, 2, 2], False, 1.0, None, False),
    ((2, 2), 3, [2, 3, 3, 2], True, 0.0, None, True),
    ((2, 2), 3, [3, 2, 2, 3], False, 0.0, 5, False),
    ((2, 1), 4, [2, 3, 2], True, 0.3, None, False),
    ((3, 3), 2, [2, 3, 2, 2, 3, 2], True, 0.0, None, False),
    ((3, 3), 2, [2, 2, 2, 2, 2, 2], False, 0.5, None, False),
]


def _ids(cases):
    return ["%d+%d_k%d_%s_s%g%s%s" % (d[0], d[1], k, "per" if p else "dir",
                                      s, "_q%d" % q if q else "",
                                      "_coll" if c else "")
            for d, k, _, p, s, q, c in cases]


@pytest.mark.parametrize("algo", ["ecl", "fcl"])
@pytest.mark.parametrize("dims,k,n,periodic,skew,nq,coll", CASES,
                         ids=_ids(CASES))
def test_deformed_apply_matches_oracle(dims, k, n, periodic, skew, nq, coll,
                                       algo):
    if algo == "fcl" and nq is not None:
        pytest.skip("the FCL inverse mass matrix on a deformed mesh takes "
                    "the operator's points (n_points = degree + 1)")
    print("case: deformed apply", algo, dims, k, n, periodic, skew, nq, coll,
          flush=True)
    P, mf, op, a = _pair(dims, k, n, periodic, skew, nq, coll,
                         seed=k + sum(dims))
    op.set_algorithm(H.ALGO_QUADRATURE if algo == "ecl" else H.ALGO_FCL)
    u = np.random.default_rng(3).standard_normal(P.n_dofs)
    src = torch.from_numpy(u).cuda()
    for t in (0.0, 0.29):
        ref = P.apply(u, t=t, fcl=(algo == "fcl"))
        dst = torch.empty_like(src)
        op.apply(dst, src, time=t)
        torch.cuda.synchronize()
        assert _rel(dst.cpu().numpy(), ref) <= 1e-12
    u32 = src.float()
    ref = P.apply(u32.double().cpu().numpy(), fcl=(algo == "fcl"))
    d32 = torch.empty_like(u32)
    op.apply(d32, u32, precision=H.FP32)
    torch.cuda.synchronize()
    assert _rel(d32.double().cpu().numpy(), ref) <= 2e-6


def test_deformed_rejects_the_collapsed_form():
    P, mf, op, a = _pair((2, 2), 3, [2, 2, 2, 2])
    assert op.algorithm() == H.ALGO_QUADRATURE
    with pytest.raises(H.HyperdealError):
        op.set_algorithm(H.ALGO_COLLAPSED)
    # the mesh cannot be deformed again while operators exist
    with pytest.raises(H.HyperdealError):
        mf.set_deformation(0.2, 2)


@pytest.mark.parametrize("dims,k,n", [((2, 2), 3, [3, 2, 2, 3]),
                                      ((3, 3), 2, [2, 2, 3, 2, 2, 2])])
def test_deformed_interpolate_norm_and_cfl(dims, k, n):
    print("case: deformed interpolate/norm/cfl", dims, k, n, flush=True)
    P, mf, op, a = _pair(dims, k, n, delta=0.12, freq=2)
    f = P.interpolate(0.1)
    g = mf.initialize_dof_vector()
    H.VectorTools.interpolate(mf, g, H.FUNC_HYPERRECTANGLE, a, 0.1)
    torch.cuda.synchronize()
    assert _rel(g.cpu().numpy(), f) <= 1e-12
    u = f + 0.01 * np.random.default_rng(4).standard_normal(P.n_dofs)
    ref = P.norm_and_error(u, 0.2)
    got = H.VectorTools.norm_and_error(mf, torch.from_numpy(u).cuda(),
                                       H.FUNC_HYPERRECTANGLE, a, 0.2)
    assert abs(got[0] - ref[0]) <= 1e-12 * ref[0]
    assert abs(got[1] - ref[1]) <= 1e-12 * ref[1]
    crit = H.compute_critical_time_step(mf, a)
    assert abs(crit - P.critical_time_step()) <= 1e-12 * crit


@pytest.mark.parametrize("algo", ["ecl", "fcl"])
def test_deformed_lsrk_steps_match_oracle(algo):
    print("case: deformed lsrk", algo, flush=True)
    P, mf, op, a = _pair((2, 2), 3, [3, 2, 2, 3], periodic=False, seed=9)
    op.set_algorithm(H.ALGO_QUADRATURE if algo == "ecl" else H.ALGO_FCL)
    u = P.interpolate() * (1 + 0.05 * np.random.default_rng(5)
                           .standard_normal(P.n_dofs))
    sol_ref, ki, ti = u.copy(), P.vector(), P.vector()
    dt = 0.1 * P.critical_time_step() / 3 ** 1.5
    for s in range(3):
        P.lsrk_step("rk45", sol_ref, ki, ti, s * dt, dt, ecl=(algo == "ecl"))
    sol = torch.from_numpy(u.copy()).cuda()
    ta, tb = torch.empty_like(sol), torch.empty_like(sol)
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    for s in range(3):
        integ.perform_time_step(op, sol, ta, tb, s * dt, dt)
    torch.cuda.synchronize()
    assert _rel(sol.cpu().numpy(), sol_ref) <= 1e-12
