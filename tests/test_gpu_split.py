"""The split form of the collapsed apply (cells too large for LDS: 3+3 at
k = 4, 5 and 2+3 / 3+2 at k = 4, 5) against the CPU oracle, through the C
ABI.

Bars:
  - FP64 apply: max |gpu - oracle| <= 1e-12 max |oracle|
  - FP32 apply: max |gpu - oracle(fp32-rounded input)| <= 2e-6 max |oracle|
  - LSRK step (rk45, FP64) through hd_lsrk_step (last stage with the free
    register as scratch) and through single stages (last stage without one):
    equal to the stage recurrence built from FP64 applies to 1e-12
"""

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
H = None

CASES = [
    ((3, 3), 5, [2, 1, 3, 2, 1, 2], [0.7, -0.4, 0.0, 0.3, -0.9, 0.55]),
    ((3, 3), 4, [1, 2, 2, 1, 3, 1], [-0.6, 0.2, 0.8, -0.35, 0.5, -0.7]),
    ((2, 3), 5, [2, 3, 1, 2, 2], [0.3, -0.8, 0.6, -0.2, 0.45]),
    ((3, 2), 4, [3, 1, 2, 2, 2], [-0.5, 0.9, -0.3, 0.65, 0.0]),
]


def setup_module(module):
    global H
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hyperdeal_amd
    H = hyperdeal_amd
    H.load_library()


def _pair(dims, k, n_cells, a, seed):
    D = sum(dims)
    r = np.random.default_rng(seed)
    hi = list(r.uniform(0.5, 2.0, D))
    P = O.Problem(dims[0], dims[1], k, n_cells, [0.0] * D, hi, a)
    mf = H.MatrixFree(dims[0], dims[1], k, n_cells, [0.0] * D, hi)
    op = H.AdvectionOperation(mf, a)
    return P, mf, op


def _rel(x, ref):
    return np.max(np.abs(x - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("dims,k,n_cells,a", CASES,
                         ids=["%d+%d_k%d" % (c[0][0], c[0][1], c[1])
                              for c in CASES])
def test_split_apply_matches_oracle(dims, k, n_cells, a, precision):
    print("case: split apply", dims, k, n_cells, precision, flush=True)
    P, mf, op = _pair(dims, k, n_cells, a, seed=k)
    assert op.algorithm() == H.ALGO_COLLAPSED
    u = np.random.default_rng(3).standard_normal(P.n_dofs)
    if precision == "fp64":
        ref = P.apply(u)
        src = torch.from_numpy(u).cuda()
        dst = torch.empty_like(src)
        op.apply(dst, src)
        tol = 1e-12
    else:
        src = torch.from_numpy(u).cuda().float()
        ref = P.apply(src.double().cpu().numpy())
        dst = torch.empty_like(src)
        op.apply(dst, src, precision=H.FP32)
        tol = 2e-6
    torch.cuda.synchronize()
    assert _rel(dst.double().cpu().numpy(), ref) <= tol


def _stage_reference(op, sol, dt, n_stages_a, n_stages_b):
    """rk45 by the recurrence: K = apply(Ti); sol += b dt K;
    Ti = sol_old + a dt K (time_integrators.templates.h:91-184)."""
    ti = sol.clone()
    k = torch.empty_like(sol)
    for s in range(len(n_stages_b)):
        op.apply(k, ti)
        old = sol.clone()
        sol += n_stages_b[s] * dt * k
        if s < len(n_stages_a):
            ti = old + n_stages_a[s] * dt * k
    return sol


def test_split_lsrk_step_matches_recurrence():
    print("case: split lsrk", flush=True)
    dims, k, n_cells, a = CASES[0]
    P, mf, op = _pair(dims, k, n_cells, a, seed=5)
    u = torch.from_numpy(
        np.random.default_rng(4).standard_normal(P.n_dofs)).cuda()
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    ca, cb = integ.ai, integ.bi
    dt = 1e-2
    ref = _stage_reference(op, u.clone(), dt, ca, cb)
    # hd_lsrk_step: the last stage's first part goes to the free register
    sol = u.clone()
    ta, tb = torch.empty_like(sol), torch.empty_like(sol)
    integ.perform_time_step(op, sol, ta, tb, 0.0, dt)
    torch.cuda.synchronize()
    assert _rel(sol.cpu().numpy(), ref.cpu().numpy()) <= 1e-12
    # single stages; the last one has no free vector (internal scratch)
    sol2 = u.clone()
    ti = [sol2.clone(), torch.empty_like(sol2)]
    for s in range(len(cb)):
        last = s == len(cb) - 1
        op.stage(ti[s % 2], None if last else ti[(s + 1) % 2], sol2, 0.0,
                 0.0 if last else ca[s] * dt, cb[s] * dt)
    torch.cuda.synchronize()
    assert _rel(sol2.cpu().numpy(), ref.cpu().numpy()) <= 1e-12


@pytest.mark.parametrize("case", [0, 1])
def test_split_fp32_stages_match_fp64(case):
    """FP32-mixed LSRK stages of 3+3 cells larger than LDS (the one-pass
    kernel whose cell result lives in a block's registers) against the FP64
    stages on the same FP32-rounded input: sol += b dt K, ti = sol + a dt K
    (middle stage) and sol += b dt K alone (last stage), within 2e-6 of
    b dt max |K| (FP32 arithmetic of K)."""
    dims, k, n_cells, a = CASES[case]
    print("case: split fp32 stages", dims, k, n_cells, flush=True)
    P, mf, op = _pair(dims, k, n_cells, a, seed=7 + case)
    r = np.random.default_rng(11)
    u32 = torch.from_numpy(r.standard_normal(P.n_dofs)).cuda().float()
    u64 = u32.double()
    sol0 = torch.from_numpy(r.standard_normal(P.n_dofs)).cuda()
    kref = torch.empty_like(u64)
    op.apply(kref, u64)
    torch.cuda.synchronize()
    kmax = kref.abs().max().item()
    ad, bd = 0.37, 0.21
    for last in (False, True):
        s64, s32 = sol0.clone(), sol0.clone()
        t64 = None if last else torch.empty_like(u64)
        t32 = None if last else torch.empty_like(u32)
        op.stage(u64, t64, s64, 0.0, 0.0 if last else ad, bd)
        op.stage(u32, t32, s32, 0.0, 0.0 if last else ad, bd,
                 precision=H.FP32)
        torch.cuda.synchronize()
        assert (s32 - s64).abs().max().item() <= 2e-6 * bd * kmax
        if not last:
            err = (t32.double() - t64).abs().max().item()
            assert err <= 2e-6 * ad * kmax + 6e-8 * t64.abs().max().item()
