"""Pins the oracle's deformed-mesh operator (no reference golden covers a
deformed mesh: the reference's DeformedCubeManifold runs, source/grid/
grid_generator.cc:110-230, have no stored output) against what the
mathematics fixes:

- the per-point geometry: the oracle's inverse Jacobian and determinant
  against central finite differences of its own mapped points (the map
  x^ + delta prod sin(f pi (x^ - lo) / (hi - lo)) (1, ..., 1) per space);
  the map fixes the box boundary and is periodic;
- deformation 0 is the Cartesian operator (which the goldens pin);
- the element-centric (ECL) and face-centric (FCL) loops, built from
  different code (cell-local faces with the cell's own normal against faces
  evaluated once from the minus side), give the same operator;
- discrete mass conservation on periodic meshes (skew 0): the JxW-weighted
  sum of the apply at the quadrature points vanishes, because every
  interior face flux enters its two cells with opposite signs;
- the convergence order of the transported product of sines on refined
  deformed meshes is that of a degree-k DG method (about k + 1).
"""

import numpy as np
import pytest

import oracle as O


def _problem(dims, k, n, periodic=True, skew=0.0, a=None, delta=0.1,
             freq=2, n_points=None, collocation=False, lo=None, hi=None):
    D = sum(dims)
    a = [0.8, -0.45, 0.3, 0.6, -0.2, 0.5][:D] if a is None else a
    lo = [0.0] * D if lo is None else lo
    hi = [1.0] * D if hi is None else hi
    P = O.Problem(dims[0], dims[1], k, n, lo, hi, a, n_points=n_points,
                  collocation=collocation, periodic_x=periodic,
                  periodic_v=periodic, skew=skew,
                  func=O.FUNC_PRODUCT_SINES, func_params=a)
    if delta:
        P.set_deformation(delta, freq)
    return P


@pytest.mark.parametrize("dims", [(1, 1), (2, 2), (3, 3), (2, 3)])
def test_geometry_matches_finite_differences(dims):
    D = sum(dims)
    P = _problem(dims, 2, [3] * D, lo=[-1.0] * D, hi=[1.0] * D, delta=0.13)
    r = np.random.default_rng(D)
    for _ in range(5):
        ci = [int(r.integers(0, 3)) for _ in range(D)]
        xi = r.uniform(0, 1, D)
        x, ji, det = P.geometry(ci, xi)
        J = np.zeros((D, D))
        eps = 1e-6
        for j in range(D):
            xp, xm = xi.copy(), xi.copy()
            xp[j] += eps
            xm[j] -= eps
            J[:, j] = (P.geometry(ci, xp)[0] - P.geometry(ci, xm)[0]) / (2 * eps)
        assert np.abs(np.linalg.inv(J) - ji).max() <= 1e-7 * np.abs(ji).max()
        assert abs(np.linalg.det(J) - det) <= 1e-8 * abs(det)
        # block diagonal over the two spaces
        assert np.all(ji[:dims[0], dims[0]:] == 0)
        assert np.all(ji[dims[0]:, :dims[0]] == 0)


def test_map_fixes_the_boundary_and_is_periodic():
    P = _problem((2, 2), 2, [4, 4, 4, 4], lo=[-1.0] * 4, hi=[1.0] * 4)
    # a point on the box boundary does not move; opposite faces match
    for xi0 in (0.0, 1.0):
        ci = [0 if xi0 == 0.0 else 3, 1, 2, 0]
        x, _, _ = P.geometry(ci, [xi0, 0.3, 0.7, 0.2])
        assert abs(x[0] - (-1.0 if xi0 == 0.0 else 1.0)) <= 1e-15
    xa, _, _ = P.geometry([0, 1, 2, 0], [0.0, 0.3, 0.7, 0.2])
    xb, _, _ = P.geometry([3, 1, 2, 0], [1.0, 0.3, 0.7, 0.2])
    assert np.allclose(xa[1:], xb[1:], atol=1e-15)
    # an interior point moves
    x, _, _ = P.geometry([1, 1, 1, 1], [0.5, 0.5, 0.5, 0.5])
    assert np.abs(x - np.array([-0.25, -0.25, -0.25, -0.25])).max() > 1e-3


def test_zero_deformation_is_cartesian():
    u = np.random.default_rng(0).standard_normal(3 * 2 * 2 * 3 * 256)
    P0 = _problem((2, 2), 3, [3, 2, 2, 3], delta=0)
    P1 = _problem((2, 2), 3, [3, 2, 2, 3], delta=0)
    P1.set_deformation(0.0, 2)
    assert np.array_equal(P0.apply(u), P1.apply(u))


CASES = [
    # dims, degree, n, periodic, skew, n_points, collocation
    ((1, 1), 3, [5, 4], True, 0.0, None, False),
    ((2, 2), 2, [3, 3, 2, 3], True, 0.0, None, False),
    ((2, 2), 3, [3, 2, 2, 3], False, 0.5, None, False),
    ((2, 2), 3, [2, 3, 3, 2], True, 0.0, None, True),
    ((2, 1), 2, [3, 2, 3], False, 1.0, None, False),
    ((3, 3), 1, [2, 3, 2, 2, 3, 2], True, 0.0, None, False),
]


@pytest.mark.parametrize("dims,k,n,periodic,skew,nq,coll", CASES)
def test_ecl_equals_fcl_on_deformed_meshes(dims, k, n, periodic, skew, nq,
                                           coll):
    P = _problem(dims, k, n, periodic=periodic, skew=skew, n_points=nq,
                 collocation=coll, lo=[-1.0] * sum(dims),
                 hi=[1.0] * sum(dims))
    u = np.random.default_rng(1).standard_normal(P.n_dofs)
    for t in (0.0, 0.3):
        e = P.apply(u, t=t)
        f = P.apply(u, t=t, fcl=True)
        assert np.abs(e - f).max() <= 1e-13 * np.abs(e).max()


@pytest.mark.parametrize("dims,k,n", [((1, 1), 3, [5, 4]),
                                      ((2, 2), 2, [3, 3, 2, 3]),
                                      ((2, 1), 3, [2, 3, 3])])
def test_mass_conservation_on_deformed_periodic_mesh(dims, k, n):
    D = sum(dims)
    P = _problem(dims, k, n, lo=[-1.0] * D, hi=[1.0] * D, delta=0.15)
    u = np.random.default_rng(2).standard_normal(P.n_dofs)
    du = P.apply(u)
    t = O.tables(k)
    N = k + 1
    xq, wq = t["qpoints"], t["qweights"]
    total, scale = 0.0, 0.0
    cell_shape = [N] * D
    for c in range(int(np.prod(n))):
        ci, r = [], c
        for d in range(D):
            ci.append(r % n[d])
            r //= n[d]
        vals = du[c * N ** D:(c + 1) * N ** D].reshape(cell_shape[::-1])
        # nodal -> Gauss points along every direction (axis D-1-d is d)
        for d in range(D):
            vals = np.moveaxis(np.tensordot(t["S"], vals,
                                            axes=([1], [D - 1 - d])), 0,
                               D - 1 - d)
        for q in np.ndindex(*([N] * D)):
            qd = q[::-1]  # qd[d]: point index of direction d
            xi = np.array([xq[i] for i in qd])
            _, _, det = P.geometry(ci, xi)
            w = np.prod([wq[i] for i in qd]) * abs(det)
            total += w * vals[q]
            scale += w * abs(vals[q])
    assert abs(total) <= 1e-12 * scale


@pytest.mark.parametrize("k", [2, 3])
def test_convergence_order_on_deformed_meshes(k):
    """1+1 periodic [0,1]^2, deformation 0.1 with frequency 2, transport of
    prod sin(2 pi (x_d - a_d t)) to T = 0.2 with rk45 at a small CFL, on
    n and 2n cells per direction: the L2 error falls by about 2^(k+1)."""
    a = [0.7, -0.4]
    errs = []
    for n in (8, 16):
        P = _problem((1, 1), k, [n, n], a=a)
        dt_crit = P.critical_time_step()
        T = 0.2
        steps = int(np.ceil(T / (0.05 * dt_crit / k ** 1.5)))
        dt = T / steps
        sol, ki, ti = P.interpolate(), P.vector(), P.vector()
        for s in range(steps):
            P.lsrk_step("rk45", sol, ki, ti, s * dt, dt)
        errs.append(P.norm_and_error(sol, T)[1])
    order = np.log2(errs[0] / errs[1])
    assert order >= k + 0.5, (errs, order)
