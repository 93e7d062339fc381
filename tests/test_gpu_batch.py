"""Batches of independent 2+2 operators, one launch per apply / LSRK stage
(through the C ABI).

Bars:
  - the batched apply and LSRK step equal the single-operator calls bit for
    bit (the batch kernel runs the same block code per operator);
  - one batched apply against the oracle at 1e-12 (FP64);
  - a hipGraph-captured LSRK step, replayed with other GPU work and
    synchronisations in between, equals eager batched steps bit for bit;
  - the autotuner leaves a valid work split and the results unchanged.
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


def _draw(seed, j):
    return O.splitmix_uniform(seed, j)


def _instances(n, sizes=(2, 3, 4), degrees=(1, 2, 3, 4, 5), seed0=20261204):
    """instance i from seed seed0 + i: cells per direction, degree and
    velocity U(-1,1)^4 drawn with SplitMix64 (the batching workload's rule,
    on small meshes)."""
    out = []
    for i in range(n):
        s = seed0 + i
        nc = sizes[int((_draw(s, 0) + 1) / 2 * len(sizes)) % len(sizes)]
        k = degrees[int((_draw(s, 1) + 1) / 2 * len(degrees)) % len(degrees)]
        a = [_draw(s, 2 + d) for d in range(4)]
        mf = H.MatrixFree(2, 2, k, [nc] * 4, [0.0] * 4, [1.0] * 4)
        op = H.AdvectionOperation(mf, a)
        out.append((mf, op, a, k, nc))
    return out


def _vectors(insts, dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return [torch.randn(mf.n_dofs(), dtype=dtype, device="cuda", generator=g)
            for mf, *_ in insts]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_apply_equals_single(precision):
    print("case: batch apply", precision, flush=True)
    prec = H.FP64 if precision == "fp64" else H.FP32
    dtype = torch.float64 if precision == "fp64" else torch.float32
    insts = _instances(12)
    batch = H.Batch([op for _, op, *_ in insts])
    src = _vectors(insts, dtype, 1)
    dst = [torch.empty_like(s) for s in src]
    batch.apply(dst, src, precision=prec)
    for (mf, op, *_), s, d in zip(insts, src, dst):
        ref = torch.empty_like(s)
        op.apply(ref, s, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(d, ref)


def test_batch_apply_matches_oracle():
    print("case: batch apply vs oracle", flush=True)
    insts = _instances(5, sizes=(2, 3))
    batch = H.Batch([op for _, op, *_ in insts])
    src = _vectors(insts, torch.float64, 2)
    dst = [torch.empty_like(s) for s in src]
    batch.apply(dst, src)
    torch.cuda.synchronize()
    for (mf, op, a, k, nc), s, d in zip(insts, src, dst):
        P = O.Problem(2, 2, k, [nc] * 4, [0.0] * 4, [1.0] * 4, a)
        ref = P.apply(s.cpu().numpy())
        got = d.cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_lsrk_step_equals_single(precision):
    print("case: batch lsrk step", precision, flush=True)
    prec = H.FP64 if precision == "fp64" else H.FP32
    dtype = torch.float64 if precision == "fp64" else torch.float32
    insts = _instances(10)
    batch = H.Batch([op for _, op, *_ in insts])
    sol = _vectors(insts, torch.float64, 3)
    sol_ref = [s.clone() for s in sol]
    ta = [torch.empty(s.numel(), dtype=dtype, device="cuda") for s in sol]
    tb = [torch.empty_like(t) for t in ta]
    dt = 1e-3
    for step in range(2):
        batch.lsrk_step("rk45", sol, ta, tb, dt, precision=prec)
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    for (mf, op, *_), s in zip(insts, sol_ref):
        a = torch.empty(s.numel(), dtype=dtype, device="cuda")
        b = torch.empty_like(a)
        for step in range(2):
            integ.perform_time_step(op, s, a, b, step * dt, dt, precision=prec)
    torch.cuda.synchronize()
    for s, r in zip(sol, sol_ref):
        assert torch.equal(s, r)


def test_batch_graph_replays_equal_eager_steps():
    print("case: batch graph replays", flush=True)
    insts = _instances(8)
    batch = H.Batch([op for _, op, *_ in insts])
    sol = _vectors(insts, torch.float64, 4)
    sol_eager = [s.clone() for s in sol]
    ta = [torch.empty(s.numel(), dtype=torch.float32, device="cuda")
          for s in sol]
    tb = [torch.empty_like(t) for t in ta]
    ea = [torch.empty_like(t) for t in ta]
    eb = [torch.empty_like(t) for t in ta]
    dt = 2e-3
    graph = batch.graph("rk45", sol, ta, tb, dt, precision=H.FP32)
    junk = torch.randn(1 << 20, device="cuda")
    for r in range(3):
        graph.launch()
        torch.cuda.synchronize()
        junk.mul_(1.0001).add_(0.5)       # other GPU work between replays
        batch.lsrk_step("rk45", sol_eager, ea, eb, dt, precision=H.FP32)
        torch.cuda.synchronize()
        for s, e in zip(sol, sol_eager):
            assert torch.equal(s, e), "replay %d" % r
    del graph


def test_batch_autotune_keeps_results():
    print("case: batch autotune", flush=True)
    insts = _instances(10)
    batch = H.Batch([op for _, op, *_ in insts])
    src = _vectors(insts, torch.float32, 5)
    dst = [torch.empty_like(s) for s in src]
    before = [torch.empty_like(s) for s in src]
    batch.apply(before, src, precision=H.FP32)
    batch.autotune(src, dst, precision=H.FP32)
    for k in sorted({k for *_, k, nc in insts}):
        assert batch.cells_per_task(k, precision=H.FP32) > 0
    batch.apply(dst, src, precision=H.FP32)
    torch.cuda.synchronize()
    for b, d in zip(before, dst):
        assert torch.equal(b, d)
    batch.set_cells_per_task(3, 1, precision=H.FP32)
    assert batch.cells_per_task(3, precision=H.FP32) >= 1


def test_batch_rejects_unsupported_operators():
    mf = H.MatrixFree(3, 3, 2, [2] * 6, [0.0] * 6, [1.0] * 6)
    op = H.AdvectionOperation(mf, [0.1] * 6)
    with pytest.raises(H.HyperdealError):
        H.Batch([op])
    mf2 = H.MatrixFree(2, 2, 2, [2] * 4, [0.0] * 4, [1.0] * 4,
                       periodic_x=False)
    op2 = H.AdvectionOperation(mf2, [0.1] * 4)
    with pytest.raises(H.HyperdealError):
        H.Batch([op2])


def test_c4_mix_graph_steps_equal_single_operator_steps():
    """The batching workload itself: 64 instances, instance i drawn from
    seed 20261204 + i (cells per direction from {4, 6, 8, 12, 16}, degree
    from {2, 3, 4, 5}, velocity U(-1, 1)^4), FP32-mixed rk45.  After the
    autotuner picks the work split, two hipGraph-replayed batch steps equal
    two single-operator steps of every instance bit for bit."""
    print("case: C4 mix, 64 instances", flush=True)
    insts = _instances(64, sizes=(4, 6, 8, 12, 16), degrees=(2, 3, 4, 5))
    assert sum(mf.n_dofs() for mf, *_ in insts) > 5e8
    batch = H.Batch([op for _, op, *_ in insts])
    sol, ta, tb = [], [], []
    for mf, op, a, k, nc in insts:
        s = mf.initialize_dof_vector()
        H.VectorTools.interpolate(mf, s, H.FUNC_PRODUCT_SINES, [0.0] * 4)
        H.VectorTools.add_noise(mf, s, 20261203, 1e-3)
        sol.append(s)
        ta.append(torch.empty(s.numel(), dtype=torch.float32, device="cuda"))
        tb.append(torch.empty_like(ta[-1]))
    for t, s in zip(ta, sol):
        t.copy_(s)
    batch.autotune(ta, tb, precision=H.FP32)
    sol_ref = [s.clone() for s in sol]
    dt = 1e-4
    graph = batch.graph("rk45", sol, ta, tb, dt, precision=H.FP32)
    for _ in range(2):
        graph.launch()
    torch.cuda.synchronize()
    del graph, ta, tb
    integ = H.LowStorageRungeKuttaIntegrator("rk45")
    for (mf, op, *_), s, r in zip(insts, sol, sol_ref):
        a = torch.empty(r.numel(), dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for step in range(2):
            integ.perform_time_step(op, r, a, b, step * dt, dt,
                                    precision=H.FP32)
        torch.cuda.synchronize()
        assert torch.equal(s, r)
