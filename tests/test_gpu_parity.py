"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars (written in each test):
  - FP64 apply, both the quadrature kernel (the reference's arithmetic step
    by step) and the collapsed kernel: max |gpu - oracle| <= 1e-12 max |oracle|
  - FP32 apply: max |gpu - oracle(fp32-rounded input)| <= 2e-6 max |oracle|
  - the reference's golden diagnostics, run end to end on the GPU
    (interpolate, CFL dt, fused LSRK steps, L2 norm/error): numdiff
    -a 1e-5 -r 1e-8, as the reference's own harness compares them
  - full size 3+3, k=3, 8^6 cells: size-independent properties (constants
    are steady, linearity, mass conservation) and oracle spot checks on
    sampled cells
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


def _rng(seed):
    return np.random.default_rng(seed)


def _pair(dims, degree, n_cells, seed, n_points=None, collocation=False,
          periodic=True, skew=0.0, bc=O.BC_INHOMOGENEOUS,
          func=O.FUNC_HYPERRECTANGLE, lo=None, hi=None, a=None):
    D = sum(dims)
    r = _rng(seed)
    a = list(r.uniform(-1, 1, D)) if a is None else a
    lo = [0.0] * D if lo is None else lo
    hi = list(r.uniform(0.5, 2.0, D)) if hi is None else hi
    P = O.Problem(dims[0], dims[1], degree, n_cells, lo, hi, a,
                  n_points=n_points, collocation=collocation,
                  periodic_x=periodic, periodic_v=periodic, skew=skew, bc=bc,
                  func=func)
    mf = H.MatrixFree(dims[0], dims[1], degree, n_cells, lo, hi,
                      periodic_x=periodic, periodic_v=periodic,
                      n_points=n_points, collocation=collocation)
    op = H.AdvectionOperation(mf, a, skew_factor=skew, boundary=bc,
                              boundary_function=func, boundary_params=a)
    return P, mf, op, a


def _rel_err(x, ref):
    return np.max(np.abs(x - ref)) / np.max(np.abs(ref))


APPLY_CASES = [
    # dims, degree, n_cells
    ((1, 1), 3, [5, 4]),
    ((1, 1), 1, [6, 5]),
    ((2, 2), 3, [3, 2, 3, 2]),
    ((2, 2), 2, [3, 3, 2, 2]),
    ((2, 2), 5, [2, 2, 2, 3]),
    ((1, 2), 3, [3, 2, 2]),
    ((2, 1), 4, [2, 3, 2]),
    ((3, 2), 3, [2, 2, 2, 2, 3]),
    ((3, 3), 3, [2, 3, 2, 2, 3, 2]),
    ((3, 3), 2, [3, 2, 2, 2, 2, 3]),
]


@pytest.mark.parametrize("a01", [(0.7, 0.3), (-0.55, -0.3), (0.0, 0.3),
                                 (0.7, -0.3), (-0.55, 0.0)])
@pytest.mark.parametrize("n0", [1, 3])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_collapsed_3d3d_row_walk(a01, n0, precision):
    """3+3, k=3: a block owns cells that are neighbours along x1 and walks
    their rows along x0 in upwind order; x0 inflow faces come from the
    previous cell, x1 inflow faces of all but one cell from the block's
    neighbour cell, both in LDS.  Rows of one and three cells, four cells
    along x1 (groups of two or four), transport to +, - and none in x0 and
    x1, against the oracle."""
    a0, a1 = a01
    print("case: row walk a0=%g a1=%g n0=%d %s" % (a0, a1, n0, precision),
          flush=True)
    a = [a0, a1, -0.8, 0.45, -0.2, 0.9]
    n_cells = [n0, 4, 3, 2, 1, 2]
    P, mf, op, a = _pair((3, 3), 3, n_cells, seed=11, a=a)
    op.set_algorithm(H.ALGO_COLLAPSED)
    r = _rng(12)
    u = P.interpolate() * (1.0 + 0.1 * r.standard_normal(P.n_dofs))
    ref = P.apply(u)
    if precision == "fp64":
        out = torch.empty(P.n_dofs, dtype=torch.float64, device="cuda")
        op.apply(out, torch.from_numpy(u).cuda())
        tol = 1e-12
    else:
        u32 = torch.from_numpy(u).cuda().float()
        ref = P.apply(u32.double().cpu().numpy())
        out = torch.empty_like(u32)
        op.apply(out, u32, precision=H.FP32)
        tol = 2e-6
    torch.cuda.synchronize()
    assert _rel_err(out.double().cpu().numpy(), ref) <= tol


@pytest.mark.parametrize("algo", ["quadrature", "collapsed"])
@pytest.mark.parametrize("dims,degree,n_cells", APPLY_CASES,
                         ids=[f"{d[0]}+{d[1]}_k{k}" for d, k, _ in
                              APPLY_CASES])
def test_apply_fp64_matches_oracle(dims, degree, n_cells, algo):
    P, mf, op, a = _pair(dims, degree, n_cells, seed=degree * 100 + sum(dims))
    op.set_algorithm(H.ALGO_QUADRATURE if algo == "quadrature"
                     else H.ALGO_COLLAPSED)
    u = _rng(1).standard_normal(P.n_dofs)
    ref = P.apply(u)
    src = torch.from_numpy(u).cuda()
    dst = torch.empty_like(src)
    op.apply(dst, src)
    torch.cuda.synchronize()
    assert _rel_err(dst.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("dims,degree,n_cells", APPLY_CASES[:8],
                         ids=[f"{d[0]}+{d[1]}_k{k}" for d, k, _ in
                              APPLY_CASES[:8]])
def test_apply_fp32_matches_oracle(dims, degree, n_cells):
    P, mf, op, a = _pair(dims, degree, n_cells, seed=7 + degree)
    u = _rng(2).standard_normal(P.n_dofs).astype(np.float32)
    ref = P.apply(u.astype(np.float64))
    src = torch.from_numpy(u).cuda()
    dst = torch.empty_like(src)
    op.apply(dst, src, precision=H.FP32)
    torch.cuda.synchronize()
    assert _rel_err(dst.cpu().numpy().astype(np.float64), ref) <= 2e-6


@pytest.mark.parametrize("skew", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("collocation", [False, True])
@pytest.mark.parametrize("algo", ["quadrature", "collapsed"])
def test_dirichlet_skew_collocation(skew, collocation, algo):
    """Dirichlet data with the skew form and collocation, through the
    quadrature kernel and through the collapsed kernel with its boundary
    correction."""
    P, mf, op, a = _pair((2, 2), 3, [3, 2, 2, 3], seed=11, periodic=False,
                         skew=skew, collocation=collocation,
                         lo=[-1.0] * 4, hi=[1.0] * 4)
    op.set_algorithm(H.ALGO_QUADRATURE if algo == "quadrature"
                     else H.ALGO_COLLAPSED)
    u = _rng(3).standard_normal(P.n_dofs)
    for t in (0.0, 0.37):
        ref = P.apply(u, t=t)
        src = torch.from_numpy(u).cuda()
        dst = torch.empty_like(src)
        op.apply(dst, src, time=t)
        torch.cuda.synchronize()
        assert _rel_err(dst.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("algo", ["quadrature", "collapsed"])
def test_homogeneous_dirichlet_3d3d(algo):
    P, mf, op, a = _pair((3, 3), 3, [2, 2, 3, 2, 2, 2], seed=12,
                         periodic=False, bc=O.BC_HOMOGENEOUS)
    op.set_algorithm(H.ALGO_QUADRATURE if algo == "quadrature"
                     else H.ALGO_COLLAPSED)
    u = _rng(4).standard_normal(P.n_dofs)
    ref = P.apply(u)
    src = torch.from_numpy(u).cuda()
    dst = torch.empty_like(src)
    op.apply(dst, src)
    torch.cuda.synchronize()
    assert _rel_err(dst.cpu().numpy(), ref) <= 1e-12


OVERINT_CASES = [
    # dims, degree, n_cells, periodic, skew
    ((1, 1), 3, [4, 3], False, 0.0),
    ((1, 1), 1, [5, 4], True, 0.5),
    ((2, 2), 3, [3, 2, 2, 3], False, 0.0),
    ((2, 2), 3, [2, 3, 2, 2], True, 1.0),
    ((2, 2), 4, [2, 2, 2, 2], False, 0.3),
    ((1, 2), 2, [3, 2, 3], False, 0.0),
    ((2, 3), 2, [2, 2, 2, 2, 2], False, 0.5),
    ((3, 3), 2, [2, 2, 2, 2, 2, 2], False, 0.0),
]


@pytest.mark.parametrize("dims,degree,n_cells,periodic,skew", OVERINT_CASES,
                         ids=[f"{d[0]}+{d[1]}_k{k}_{'per' if p else 'dir'}"
                              for d, k, _, p, _s in OVERINT_CASES])
def test_quadrature_overintegration_matches_oracle(dims, degree, n_cells,
                                                   periodic, skew):
    """The quadrature kernel with k+2 Gauss points per direction (the
    reference's over-integrated runs, e.g. tests/advection/adv_2D_2D_k3_q5):
    FP64 apply at 1e-12 and FP32 at 2e-6 against the oracle, inhomogeneous
    Dirichlet or periodic, and one fused LSRK stage."""
    print("case: overintegration %s k=%d %s skew=%g" %
          (dims, degree, "periodic" if periodic else "dirichlet", skew),
          flush=True)
    P, mf, op, a = _pair(dims, degree, n_cells, seed=31 + degree,
                         n_points=degree + 2, periodic=periodic, skew=skew,
                         lo=[-1.0] * sum(dims), hi=[1.0] * sum(dims))
    op.set_algorithm(H.ALGO_QUADRATURE)
    u = _rng(5).standard_normal(P.n_dofs)
    t = 0.21
    ref = P.apply(u, t=t)
    src = torch.from_numpy(u).cuda()
    dst = torch.empty_like(src)
    op.apply(dst, src, time=t)
    u32 = src.float()
    ref32 = P.apply(u32.double().cpu().numpy(), t=t)
    dst32 = torch.empty_like(u32)
    op.apply(dst32, u32, time=t, precision=H.FP32)
    torch.cuda.synchronize()
    assert _rel_err(dst.cpu().numpy(), ref) <= 1e-12
    assert _rel_err(dst32.double().cpu().numpy(), ref32) <= 2e-6
    # fused stage: sol += b dt K, ti = sol_old + a dt K with K = apply(src)
    sol0 = _rng(6).standard_normal(P.n_dofs)
    sol = torch.from_numpy(sol0).cuda()
    ti = torch.empty_like(sol)
    op.stage(src, ti, sol, t, 0.3, 0.7)
    torch.cuda.synchronize()
    k = ref
    np.testing.assert_allclose(sol.cpu().numpy(), sol0 + 0.7 * k,
                               rtol=0, atol=1e-12 * np.max(np.abs(k)) + 1e-14)
    np.testing.assert_allclose(ti.cpu().numpy(), sol0 + 0.3 * k,
                               rtol=0, atol=1e-12 * np.max(np.abs(k)) + 1e-14)


@pytest.mark.parametrize("dims,degree,n_points", [((3, 3), 5, 7),
                                                 ((3, 3), 4, 7)])
def test_quadrature_unsupported_points_rejected(dims, degree, n_points):
    """Inhomogeneous Dirichlet data: the collapsed kernel's boundary
    correction takes at most 7776 face quadrature points (n_points^(D-1)),
    the quadrature kernel k+1 and k+2 points (k+2 in 3+3 only up to k=2):
    other sizes are an error before any launch, not a silent fallback."""
    D = sum(dims)
    mf = H.MatrixFree(dims[0], dims[1], degree, [2] * D, [0.0] * D,
                      [1.0] * D, periodic_x=False, periodic_v=False,
                      n_points=n_points)
    op = H.AdvectionOperation(mf, [0.3] * D)
    src = torch.zeros(mf.n_dofs(), dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    with pytest.raises(H.HyperdealError, match="no kernel"):
        op.apply(dst, src)


@pytest.mark.parametrize("n_points,collocation", [(5, False), (6, False),
                                                  (4, True)])
def test_collapsed_overintegration_and_collocation(n_points, collocation):
    P, mf, op, a = _pair((2, 2), 3, [3, 2, 2, 3], seed=13, n_points=n_points,
                         collocation=collocation)
    assert op.algorithm() == H.ALGO_COLLAPSED
    u = _rng(5).standard_normal(P.n_dofs)
    ref = P.apply(u)
    src = torch.from_numpy(u).cuda()
    dst = torch.empty_like(src)
    op.apply(dst, src)
    torch.cuda.synchronize()
    assert _rel_err(dst.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("algo", ["quadrature", "collapsed"])
@pytest.mark.parametrize("rk", ["rk33", "rk45", "rk47", "rk59"])
def test_lsrk_step_matches_oracle(rk, algo):
    P, mf, op, a = _pair((2, 2), 3, [3, 2, 2, 3], seed=14)
    op.set_algorithm(H.ALGO_QUADRATURE if algo == "quadrature"
                     else H.ALGO_COLLAPSED)
    u = _rng(6).standard_normal(P.n_dofs)
    sol_ref, ki, ti = u.copy(), P.vector(), P.vector()
    dt = 0.01
    for step in range(3):
        P.lsrk_step(rk, sol_ref, ki, ti, step * dt, dt)
    sol = torch.from_numpy(u.copy()).cuda()
    ta, tb = torch.empty_like(sol), torch.empty_like(sol)
    integ = H.LowStorageRungeKuttaIntegrator(rk)
    for step in range(3):
        integ.perform_time_step(op, sol, ta, tb, step * dt, dt)
    torch.cuda.synchronize()
    assert _rel_err(sol.cpu().numpy(), sol_ref) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lsrk_3d3d_row_walk(precision):
    """Fused LSRK stages of the 3+3, k=3 row-walk kernel (groups of cells
    along x1, stage with and without the next stage input) against the
    oracle's rk45 steps.  FP32-mixed: the stage input is float, the solution
    and the update are double; the bar is that of float rounding carried by
    three steps."""
    print("case: lsrk 3+3 row walk", precision, flush=True)
    a = [0.6, -0.45, 0.8, -0.3, 0.55, -0.9]
    P, mf, op, a = _pair((3, 3), 3, [3, 4, 2, 2, 1, 2], seed=16, a=a)
    op.set_algorithm(H.ALGO_COLLAPSED)
    u = P.interpolate() * (1.0 + 0.1 * _rng(7).standard_normal(P.n_dofs))
    sol_ref, ki, ti = u.copy(), P.vector(), P.vector()
    dt = 0.1 * H.compute_critical_time_step(mf, a) / 3 ** 1.5
    for step in range(3):
        P.lsrk_step("rk45", sol_ref, ki, ti, step * dt, dt)
    prec, dtype, tol = ((H.FP64, torch.float64, 1e-12) if precision == "fp64"
                        else (H.FP32, torch.float32, 1e-6))
    sol = torch.from_numpy(u.copy()).cuda()
    ta = torch.empty(P.n_dofs, dtype=dtype, device="cuda")
    tb = torch.empty_like(ta)
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    for step in range(3):
        integ.perform_time_step(op, sol, ta, tb, step * dt, dt,
                                precision=prec)
    torch.cuda.synchronize()
    assert _rel_err(sol.cpu().numpy(), sol_ref) <= tol


def test_lsrk_fp32_mixed_tracks_fp64():
    P, mf, op, a = _pair((2, 2), 3, [4, 4, 4, 4], seed=15,
                         func=O.FUNC_PRODUCT_SINES, lo=[0.0] * 4,
                         hi=[1.0] * 4)
    f0 = torch.empty(mf.n_dofs(), dtype=torch.float64, device="cuda")
    H.VectorTools.interpolate(mf, f0, H.FUNC_PRODUCT_SINES, [0.0] * 4)
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    dt = 0.1 * H.compute_critical_time_step(mf, a) / 3 ** 1.5
    sols = {}
    for prec, dtype in ((H.FP64, torch.float64), (H.FP32, torch.float32)):
        sol = f0.clone()
        ta = torch.empty(mf.n_dofs(), dtype=dtype, device="cuda")
        tb = torch.empty_like(ta)
        for s in range(10):
            integ.perform_time_step(op, sol, ta, tb, s * dt, dt, precision=prec)
        sols[prec] = sol
    torch.cuda.synchronize()
    diff = (sols[H.FP64] - sols[H.FP32]).cpu().numpy()
    # L2-type error of the FP32-mixed run against the FP64 run, relative
    assert np.sqrt(np.mean(diff ** 2)) / \
        np.sqrt(np.mean(sols[H.FP64].cpu().numpy() ** 2)) <= 1e-6


def _gpu_run_advection(c, algo=None):
    """examples/advection end to end on the GPU path: interpolate f0 at the
    nodes, dt rule (application.h:363-396), time loop (time_loop.cc:33-67)
    with fused LSRK steps, diagnostics rows (application.h:464-506); algo:
    the algorithm forced on the operator (None: the automatic choice)."""
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
    if algo is not None:
        op.set_algorithm(algo)
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
    t, step = t0 + dt, 1
    while t <= T * 1.0000000000001:
        integ.perform_time_step(op, sol, ta, tb, t - dt, dt)
        diag(t)
        t += dt
        step += 1
    return np.array(rows)


GOLD = AC.cases()


@pytest.mark.parametrize("c", GOLD, ids=[c["name"] for c in GOLD])
def test_reference_golden_end_to_end_on_gpu(c):
    """The automatic choice (the collapsed kernel, with the boundary
    correction for Dirichlet cases) and, for Dirichlet cases, the
    quadrature kernel too."""
    print("case: golden %s" % c["name"], flush=True)
    algos = [None]
    if not (c["periodic_x"] and c["periodic_v"]):
        algos.append(H.ALGO_QUADRATURE)
    for algo in algos:
        rows = _gpu_run_advection(c, algo)
        ours = AC.parse_rows(AC.format_rows(rows))
        assert ours.shape == c["golden"].shape
        assert AC.numdiff_ok(ours, c["golden"]), algo


# --------------------------------------------------------------------------
# full size: 3+3, k=3, 8^6 cells (1,073,741,824 DoF)
# --------------------------------------------------------------------------
C2 = dict(dims=(3, 3), degree=3, n_cells=[8] * 6)


@pytest.fixture(scope="module")
def c2():
    mf = H.MatrixFree(3, 3, 3, C2["n_cells"], [0.0] * 6, [1.0] * 6)
    a = [O.splitmix_uniform(20261202, i) for i in range(6)]
    op = H.AdvectionOperation(mf, a)
    return mf, op, a


def test_full_size_constant_is_steady(c2):
    mf, op, a = c2
    print("case: 3+3 k=3 8^6 constant input", flush=True)
    src = torch.ones(mf.n_dofs(), dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    op.apply(dst, src)
    torch.cuda.synchronize()
    scale = max(abs(x) for x in a) * 8 * 16   # |a|/h times the matrix norm
    assert dst.abs().max().item() <= 1e-12 * scale
    del src, dst


def test_full_size_linearity_and_mass(c2):
    mf, op, a = c2
    print("case: 3+3 k=3 8^6 linearity", flush=True)
    u = torch.empty(mf.n_dofs(), dtype=torch.float64, device="cuda")
    H.VectorTools.interpolate(mf, u, H.FUNC_PRODUCT_SINES, [0.0] * 6)
    H.VectorTools.add_noise(mf, u, 20261203, 1e-3)
    v = torch.empty_like(u)
    H.VectorTools.interpolate(mf, v, H.FUNC_CONSTANT_ONE)
    au, av = torch.empty_like(u), torch.empty_like(u)
    op.apply(au, u)
    op.apply(av, v)
    w = u + 2.0 * v
    aw = torch.empty_like(u)
    op.apply(aw, w)
    torch.cuda.synchronize()
    lin = (aw - au - 2.0 * av).abs().max().item()
    assert lin <= 1e-12 * au.abs().max().item()
    # mass: the cell means of the apply sum to zero (periodic transport);
    # with GL nodal values the cell mean is the GL-weighted sum
    del v, av, w, aw
    torch.cuda.empty_cache()
    x, wq = np.polynomial.legendre.leggauss(4)
    t = O.tables(3)
    # exact integration of the nodal polynomial: S^T w at Gauss points
    w1 = torch.tensor(t["S"].T @ (wq / 2), dtype=torch.float64, device="cuda")
    wcell = w1
    for _ in range(5):
        wcell = torch.outer(w1, wcell).reshape(-1)
    mass_rate = (au.view(-1, 4096) @ wcell).sum().item()
    mass = (u.view(-1, 4096) @ wcell).sum().item()
    assert abs(mass_rate) <= 1e-10 * au.abs().max().item()
    assert abs(mass) < 1.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_full_size_spot_check_against_oracle(c2, precision):
    mf, op, a = c2
    print("case: 3+3 k=3 8^6 oracle spot check", precision, flush=True)
    u = torch.empty(mf.n_dofs(), dtype=torch.float64, device="cuda")
    H.VectorTools.interpolate(mf, u, H.FUNC_PRODUCT_SINES, [0.0] * 6)
    H.VectorTools.add_noise(mf, u, 20261203, 1e-3)
    if precision == "fp32":
        u32 = u.float()
        out = torch.empty_like(u32)
        op.apply(out, u32, precision=H.FP32)
        host_u = u32.double().cpu().numpy()
        # FP32 arithmetic: each output sums ~30 products of size up to
        # |a|/h * max|A| * |u| (~8 * 20 here) into O(1) results, so float
        # rounding (6e-8) is amplified by that cancellation
        tol = 1e-5
        del u
    else:
        out = torch.empty_like(u)
        op.apply(out, u)
        host_u = u.cpu().numpy()
        tol = 1e-12
    torch.cuda.synchronize()
    P = O.Problem(3, 3, 3, C2["n_cells"], [0.0] * 6, [1.0] * 6, a)
    r = _rng(99)
    cells = np.unique(np.concatenate([
        r.integers(0, 8 ** 6, 56),
        # corners of the periodic box (wrap-around neighbours)
        [0, 7, 8 ** 6 - 1, 8 ** 5 * 7, 8 ** 3 * 7 + 7, 8 ** 6 - 8,
         8 ** 4 * 7, 8 ** 2 * 7]]))
    ref = P.apply_cells(host_u, cells)
    got = out.view(-1, 4096)[torch.from_numpy(cells).cuda()].cpu().numpy()
    exp = ref.reshape(-1, 4096)[cells]
    assert np.max(np.abs(got - exp)) <= tol * np.max(np.abs(exp))


def test_c1_plain_method_ten_steps():
    """The plain CPU-runnable configuration on the GPU: 2+2 periodic
    [0,1]^4, 8^4 cells, k=3 (1,048,576 DoF), a = (1, 0.5, 0.25, 0.125),
    f0 = prod sin(2 pi x_i) at the Gauss-Lobatto nodes, rk45 with dt from
    CFL 0.1 (0.1 dt_crit / k^1.5), 10 steps.  FP64 against the oracle's 10
    steps at 1e-12; both against the exact shifted solution f0(x - a t)
    (the L2 error of the k=3 discretisation on h = 1/8, the same for both
    to 1e-9 relative, and well below that of transport in the wrong
    direction); FP32-mixed within 1e-6 relative L2 of the FP64 run."""
    print("case: C1 8^4 k=3 rk45 10 steps", flush=True)
    a = [1.0, 0.5, 0.25, 0.125]
    n = [8] * 4
    P = O.Problem(2, 2, 3, n, [0.0] * 4, [1.0] * 4, a,
                  func=O.FUNC_PRODUCT_SINES, func_params=a)
    mf = H.MatrixFree(2, 2, 3, n, [0.0] * 4, [1.0] * 4)
    op = H.AdvectionOperation(mf, a)
    assert op.algorithm() == H.ALGO_COLLAPSED
    crit = H.compute_critical_time_step(mf, a)
    assert abs(crit - 0.125) <= 1e-15
    assert abs(P.critical_time_step() - crit) <= 1e-15
    dt = 0.1 * crit / 3 ** 1.5
    f0 = P.interpolate()
    sol_ref, ki, ti = f0.copy(), P.vector(), P.vector()
    for s in range(10):
        P.lsrk_step("rk45", sol_ref, ki, ti, s * dt, dt)
    T = 10 * dt
    runs = {}
    for prec, dtype in ((H.FP64, torch.float64), (H.FP32, torch.float32)):
        sol = mf.initialize_dof_vector()
        H.VectorTools.interpolate(mf, sol, H.FUNC_PRODUCT_SINES, a)
        ta = torch.empty(mf.n_dofs(), dtype=dtype, device="cuda")
        tb = torch.empty_like(ta)
        integ = H.LowStorageRungeKuttaIntegrator("rk45")
        for s in range(10):
            integ.perform_time_step(op, sol, ta, tb, s * dt, dt,
                                    precision=prec)
        torch.cuda.synchronize()
        runs[prec] = sol
    got = runs[H.FP64].cpu().numpy()
    assert _rel_err(got, sol_ref) <= 1e-12
    norm, err = H.VectorTools.norm_and_error(mf, runs[H.FP64],
                                             H.FUNC_PRODUCT_SINES, a, T)
    rnorm, rerr = P.norm_and_error(sol_ref, T)
    assert abs(err - rerr) <= 1e-9 * rerr and abs(norm - rnorm) <= 1e-12
    # the exact solution's L2 norm is 2^-2 (prod of four sin^2 means); the
    # error is the k=3 discretisation's (transport the wrong way would give
    # about 0.09)
    assert abs(norm - 0.25) <= 5e-3 and err <= 5e-3
    d = (runs[H.FP64] - runs[H.FP32]).cpu().numpy()
    assert np.sqrt(np.sum(d * d) / np.sum(got * got)) <= 1e-6
