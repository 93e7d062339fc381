"""The exact start of the Vlasov-Poisson potential solve (fast
diagonalisation of the periodic 2D SIPG system, hd_vp_exact_start) through
the C ABI.

Bars:
  - the potential with the exact start against the oracle's CG solve (both
    at 1e-12 relative residual, mean removed): 1e-9 of its size, the bar of
    the default start (test_gpu_vlasov.py); x meshes with n0 != n1 and
    h0 != h1 catch a transposed direction
  - the start itself solves the system: relative residual of the start
    <= 1e-11, at most one CG iteration per solve
  - along an LSRK run the field energies equal those of the default start
    to 1e-9 relative
"""

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
H = None
PI = np.pi


def setup_module(module):
    global H
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import hyperdeal_amd
    H = hyperdeal_amd


CASES = [
    (3, [8, 4, 4, 4], [4 * PI, 2 * PI]),
    (2, [6, 10, 2, 2], [4 * PI, 5 * PI]),
    (1, [5, 7, 2, 2], [2 * PI, 3 * PI]),
]


@pytest.mark.parametrize("degree,n_cells,lx", CASES,
                         ids=["k3_8x4", "k2_6x10", "k1_5x7"])
def test_exact_start_matches_oracle(degree, n_cells, lx):
    print("case: exact start", degree, n_cells, flush=True)
    lo, hi = [0, 0, -6, -6], [lx[0], lx[1], 6, 6]
    P = O.VlasovPoisson(2, degree, n_cells, lo, hi, [1, 1, 6, 6], eps=0.01,
                        kx=2 * PI / lx[0], n_perturbed=2)
    mf = H.MatrixFree(2, 2, degree, n_cells, lo, hi)
    op = H.VlasovPoissonOperation(mf, poisson_tol=1e-12)
    op.exact_start()
    r = np.random.default_rng(11)
    f = P.interpolate() * (1.0 + 0.05 * r.standard_normal(P.n_dofs))
    op.update_field(torch.from_numpy(f).cuda())
    phi_gpu = op.potential()
    phi_ref, _ = P.poisson_solve(P.rhs(P.density(f)), rel_tol=1e-12)
    a = phi_gpu - phi_gpu.mean()
    e = phi_ref - phi_ref.mean()
    assert np.max(np.abs(a - e)) <= 1e-9 * np.max(np.abs(e))
    iters, solves, res = op.solver_stats()
    r0, _, _, guessed = op.solver_start()
    print("  iterations %d, start residual %.2e" % (iters, r0), flush=True)
    assert solves == 1 and guessed == 1
    assert r0 <= 1e-11 and iters <= 1 and res <= 1e-12


def test_exact_start_along_lsrk_run():
    print("case: exact start lsrk", flush=True)
    n_cells = [8, 8, 4, 4]
    lo, hi = [0, 0, -6, -6], [4 * PI, 4 * PI, 6, 6]
    energies, its = {}, {}
    for exact in (False, True):
        mf = H.MatrixFree(2, 2, 3, n_cells, lo, hi)
        op = H.VlasovPoissonOperation(mf, poisson_tol=1e-10)
        if exact:
            op.exact_start()
        f = mf.initialize_dof_vector()
        H.VectorTools.interpolate(mf, f, H.FUNC_LANDAU, [0.05, 0.5, 2.0])
        ta, tb = torch.empty_like(f), torch.empty_like(f)
        integ = H.LowStorageRungeKuttaIntegrator("rk45")
        en = []
        for s in range(4):
            integ.perform_time_step(op, f, ta, tb, s * 0.05, 0.05)
            en.append(sum(op.field_energy()))
        energies[exact] = np.array(en)
        it, solves, _ = op.solver_stats()
        its[exact] = it / solves
    print("  CG iterations per solve: previous phi %.1f, exact start %.1f"
          % (its[False], its[True]), flush=True)
    assert np.allclose(energies[True], energies[False], rtol=1e-9, atol=0)
    assert its[True] <= 1.0 < its[False]


def test_exact_start_rejects_unsupported():
    mf = H.MatrixFree(1, 1, 3, [8, 8], [0, -6], [4 * PI, 6])
    op = H.VlasovPoissonOperation(mf)
    with pytest.raises(H.HyperdealError):
        op.exact_start()
