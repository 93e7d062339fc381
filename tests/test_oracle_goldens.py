"""Pins the CPU oracle to every golden file the reference ships for the path.

- examples/advection/tests/*.out: the advection application end to end
  (interpolate f0, CFL dt, LSRK, ECL or FCL apply, L2 norm/error rows) for
  1+1 and 2+2, k=3: periodic and Dirichlet, skew 0 and 0.5, collocation,
  over-integration (N_POINTS=5) and rk33/45/47/59.  Compared on the printed
  values with the reference's numdiff tolerance (-a 1e-5 -r 1e-8), and more
  strictly on the norm/error columns.
- tests/time_discretization/time_integrators_02: y' = y sin^2 t, rk45,
  dt=0.1, 100 steps -> 118.127 (exact exp((10 - sin10 cos10)/2)).
- tests/time_discretization/time_integrators_01: stage counts 3/5/7/9.
- tests/vector_tools/vector_tools_01: interpolate + norm_and_error, 1+1,
  k=3, 2^r cells per direction.
"""

import os
import re

import numpy as np
import pytest

import oracle as O
from tests.golden import advection_cases as AC

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = AC.cases()


def _problem(c):
    D = c["dim_x"] + c["dim_v"]
    lo, hi = AC.HYPERRECTANGLE_BOX
    return O.Problem(c["dim_x"], c["dim_v"], c["degree"], c["n_cells"],
                     [lo] * D, [hi] * D, AC.HYPERRECTANGLE_VELOCITY[:D],
                     n_points=c["n_points"], collocation=c["collocation"],
                     periodic_x=c["periodic_x"], periodic_v=c["periodic_v"],
                     skew=c["skew"], bc=O.BC_INHOMOGENEOUS,
                     func=O.FUNC_HYPERRECTANGLE)


def test_all_reference_cases_found():
    assert len(CASES) == 19


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_advection_golden(c):
    P = _problem(c)
    rows, dt = P.run_advection(rk=c["rk"], time_step=c["time_step"],
                               start_time=c["start_time"],
                               final_time=c["final_time"], cfl=c["cfl"],
                               diag_tick=c["tick"], ecl=c["use_ecl"])
    ours = AC.parse_rows(AC.format_rows(rows))
    gold = c["golden"]
    assert ours.shape == gold.shape
    assert AC.numdiff_ok(ours, gold)
    # stricter than numdiff: norm and error agree to the printed digits
    for col in (1, 2):
        d = np.abs(ours[:, col] - gold[:, col])
        assert np.all((d <= 1e-14) | (d <= 2e-10 * np.abs(gold[:, col])))


def test_time_integrators_02_golden():
    txt = open(os.path.join(GOLDEN, "time_integrators",
                            "time_integrators_02.mpirun=1.output")).read()
    vals = [float(v) for v in re.findall(r"= ([0-9.]+)", txt)]
    y = O.lsrk_ode("rk45", 0.1, 100, 1.0)
    assert all(abs(v - 118.127) < 1e-12 for v in vals)
    assert float("%g" % y) == 118.127
    exact = np.exp((10.0 - np.sin(10.0) * np.cos(10.0)) / 2.0)
    assert abs(y - exact) / exact < 2e-6


def test_time_integrators_01_stage_counts():
    txt = open(os.path.join(GOLDEN, "time_integrators",
                            "time_integrators_01.output")).read()
    counts = [int(v) for v in re.findall(r"DEAL::(\d+)", txt)]
    lib = O.load()
    ours = [lib.hdo_lsrk_n_stages(k.encode())
            for k in ("rk33", "rk45", "rk47", "rk59")]
    assert ours == counts


@pytest.mark.parametrize("rk,order", [("rk33", 3), ("rk45", 4), ("rk47", 4),
                                      ("rk59", 5)])
def test_lsrk_convergence_order(rk, order):
    """Halving dt shrinks the ODE error by ~2^order (closed-form check that a
    mistyped coefficient would fail)."""
    exact = np.exp((10.0 - np.sin(10.0) * np.cos(10.0)) / 2.0)
    e1 = abs(O.lsrk_ode(rk, 0.2, 50) - exact)
    e2 = abs(O.lsrk_ode(rk, 0.1, 100) - exact)
    assert np.log2(e1 / e2) > order - 0.5


def test_vector_tools_01_golden():
    txt = open(os.path.join(GOLDEN, "vector_tools",
                            "vector_tools_01.mpirun=1.output")).read()
    vals = [tuple(float(x) for x in l.split("::")[1].split())
            for l in txt.splitlines() if "::" in l]
    ours = []
    for refs in range(2, 6):
        n = 2 ** refs
        for f in (O.FUNC_ZERO, O.FUNC_CONSTANT_ONE, O.FUNC_SIN_COS):
            P = O.Problem(1, 1, 3, [n, n], [0, 0], [1, 1], [0, 0], func=f)
            ours.append(tuple(P.norm_and_error(P.interpolate())))
    assert len(ours) == len(vals)
    for (n0, e0), (n1, e1) in zip(ours, vals):
        assert float("%g" % n0) == pytest.approx(n1, rel=1e-5, abs=1e-12)
        assert abs(e0 - e1) <= max(1e-15, 6e-6 * abs(e1))


def _vp_golden():
    rows = []
    for l in open(os.path.join(GOLDEN, "vlasov_poisson",
                               "vp_2D_2D_k3.hyperrectangle_01.out")):
        if not l.startswith("#"):
            rows.append([float(x) for x in l.split()])
    return np.array(rows)


def test_vlasov_poisson_golden():
    """examples/vlasov_poisson/tests/vp_2D_2D_k3.hyperrectangle_01: Landau
    damping on [0,4pi]^2 x [-6,6]^2, 4^4 cells, k=3, CFL 0.05, T=0.5; rows
    (t, field energy x2, mass, ||f||, kinetic energy, momentum x3).  The
    reference prints t with 3 decimals (application.h:632-642)."""
    P = O.VlasovPoisson(2, 3, [4, 4, 4, 4], [0, 0, -6, -6],
                        [4 * np.pi, 4 * np.pi, 6, 6], [1, 1, 6, 6],
                        eps=0.01, kx=0.5, n_perturbed=1)
    rows, dt, it = P.run(final_time=0.5, cfl=0.05, diag_tick=0.1,
                         poisson_tol=1e-12)
    gold = _vp_golden()
    rows[:, 0] = [float("%.3f" % t) for t in rows[:, 0]]
    assert rows.shape == gold.shape
    assert AC.numdiff_ok(rows, gold)
    # the damped field energy and the invariants to ~1e-9
    assert np.allclose(rows[:, 1], gold[:, 1], rtol=1e-8, atol=1e-12)
    assert np.allclose(rows[:, 3:6], gold[:, 3:6], rtol=1e-10, atol=0)


def test_vlasov_poisson_pieces():
    """rho, Poisson and grad phi against the closed form for the Landau
    density: -lap phi = -(rho - mean) with rho = 1 + eps cos(k x0) gives
    phi = -eps cos(k x0) / k^2 and d phi/dx0 = eps sin(k x0) / k."""
    P = O.VlasovPoisson(2, 3, [8, 8, 4, 4], [0, 0, -6, -6],
                        [4 * np.pi, 4 * np.pi, 6, 6], [1, 1, 6, 6])
    f = P.interpolate()
    rho = P.density(f)
    b = P.rhs(rho)
    assert abs(b.sum()) < 1e-12
    phi, it = P.poisson_solve(b, rel_tol=1e-12)
    assert np.linalg.norm(b - P.poisson_apply(phi)) <= 1e-11 * np.linalg.norm(b)
    P.set_potential(phi)
    E = P.efield().reshape(64, 16, 2)
    t = O.tables(3)
    h = 4 * np.pi / 8
    x0 = np.array([(c % 8 + t["qpoints"][q % 4]) * h for c in range(64)
                   for q in range(16)]).reshape(64, 16)
    # the v-integral of the interpolated Maxwellian is not exactly 1; scale
    mass_v = rho.mean()
    exact = 0.01 * mass_v * np.sin(0.5 * x0) / 0.5
    assert np.max(np.abs(E[:, :, 0] - exact)) < 2e-3 * np.max(np.abs(exact))
    assert np.max(np.abs(E[:, :, 1])) < 1e-12
