"""Pins the oracle's building blocks against what the mathematics fixes.

- 1D tables against numpy's Gauss-Legendre rule and closed forms
  (Gauss-Lobatto nodes for k=3, polynomial reproduction, Sinv S = I).
- The apply against an independent numpy construction of the same DG
  operator for constant velocity on a Cartesian mesh: the exact-integration
  (collapsed) form sum_d (a_d/h_d) Mhat^{-1} [Khat u - e_out u_out + e_in
  u_nb], built from numpy polynomials, not from the oracle's tables.
- Invariants: apply(const) = 0 (periodic), mass conservation, upwind
  dissipation, skew 0.5 == skew 0 for constant a, ECL == FCL.
"""

import numpy as np
import numpy.polynomial.legendre as npl
import numpy.polynomial.polynomial as npp
import pytest

import oracle as O


@pytest.mark.parametrize("nq", [2, 3, 4, 5, 6, 7])
def test_gauss_rule_matches_numpy(nq):
    t = O.tables(nq - 1, nq)
    x, w = npl.leggauss(nq)
    assert np.allclose(t["qpoints"], (x + 1) / 2, atol=1e-15, rtol=0)
    assert np.allclose(t["qweights"], w / 2, atol=1e-15, rtol=0)


def test_gauss_lobatto_k3_closed_form():
    t = O.tables(3)
    s = 1 / np.sqrt(5)
    assert np.allclose(t["nodes"], [0, (1 - s) / 2, (1 + s) / 2, 1],
                       atol=1e-15)
    tc = O.tables(3, collocation=True)
    assert np.allclose(tc["qweights"], [1 / 12, 5 / 12, 5 / 12, 1 / 12],
                       atol=1e-15)


@pytest.mark.parametrize("k,nq", [(1, 2), (2, 3), (3, 4), (3, 5), (4, 5),
                                  (5, 6)])
def test_tables_reproduce_polynomials(k, nq):
    t = O.tables(k, nq)
    xd, xq = t["nodes"], t["qpoints"]
    for p in range(k + 1):
        # S interpolates degree <= k exactly
        assert np.allclose(t["S"] @ xd ** p, xq ** p, atol=1e-14)
    for p in range(nq):
        # D differentiates degree <= nq-1 exactly at the quadrature points
        dp = p * xq ** (p - 1) if p > 0 else np.zeros(nq)
        assert np.allclose(t["D"] @ xq ** p, dp, atol=1e-12)
        # face vectors evaluate at 0 and 1
        assert np.allclose(t["F"] @ xq ** p, [0.0 ** p, 1.0], atol=1e-14)
    assert np.allclose(t["Sinv"] @ t["S"], np.eye(k + 1), atol=1e-13)


def _lagrange_polys(nodes):
    polys = []
    for j in range(len(nodes)):
        c = np.array([1.0])
        for m, xm in enumerate(nodes):
            if m != j:
                c = npp.polymul(c, [-xm, 1.0]) / (nodes[j] - xm)
        polys.append(c)
    return polys


def _collapsed_1d(k):
    """Mhat, Khat (Khat_ij = int l_i' l_j) on [0,1] for GL nodes, exact."""
    x, _ = npl.leggauss(k + 1)
    s = 1 / np.sqrt(5) if k == 3 else None
    # Gauss-Lobatto nodes: endpoints + roots of P_k'
    dPk = npl.legder([0] * k + [1])
    inner = np.sort(npl.legroots(dPk)) if k > 1 else np.array([])
    nodes = (np.concatenate([[-1.0], inner, [1.0]]) + 1) / 2
    if s is not None:
        assert np.allclose(nodes, [0, (1 - s) / 2, (1 + s) / 2, 1])
    L = _lagrange_polys(nodes)
    n = k + 1

    def integ(c):
        ci = npp.polyint(c)
        return npp.polyval(1.0, ci) - npp.polyval(0.0, ci)

    M = np.array([[integ(npp.polymul(L[i], L[j])) for j in range(n)]
                  for i in range(n)])
    K = np.array([[integ(npp.polymul(npp.polyder(L[i]), L[j]))
                   for j in range(n)] for i in range(n)])
    return M, K


def _apply_collapsed(u, dims, n_cells, h, a, k):
    """numpy reference of the periodic constant-velocity DG operator."""
    D = len(n_cells)
    n = k + 1
    M, K = _collapsed_1d(k)
    Minv = np.linalg.inv(M)
    # axes: cells (slowest .. fastest reversed) then nodes; build an array
    # with shape (c_{D-1},...,c_0, i_{D-1},...,i_0) from the flat layout
    U = u.reshape(list(reversed(n_cells)) + [n] * D)
    out = np.zeros_like(U)
    for d in range(D):
        cax = D - 1 - d          # cell axis of direction d
        nax = D + (D - 1 - d)    # node axis of direction d
        Ad = a[d] / h[d]
        Ud = np.moveaxis(U, nax, -1)
        r = np.einsum("ij,...j->...i", K, Ud)
        if a[d] >= 0:
            # outflow at 1 (own last node), inflow at 0 from the left
            # neighbour's last node
            r[..., n - 1] -= Ud[..., n - 1]
            nb = np.roll(Ud[..., n - 1], 1, axis=cax)
            r[..., 0] += nb
        else:
            r[..., 0] += Ud[..., 0]
            nb = np.roll(Ud[..., 0], -1, axis=cax)
            r[..., n - 1] -= nb
        r = Ad * np.einsum("ij,...j->...i", Minv, r)
        out += np.moveaxis(r, -1, nax)
    return out.reshape(-1)


@pytest.mark.parametrize("dims,n_cells", [((1, 1), [5, 3]),
                                          ((2, 2), [3, 2, 3, 2]),
                                          ((1, 2), [3, 2, 2]),
                                          ((3, 3), [2, 2, 2, 2, 2, 2])])
def test_apply_matches_numpy_collapsed_operator(dims, n_cells):
    rng = np.random.default_rng(7)
    D = sum(dims)
    a = rng.uniform(-1, 1, D)
    lo = np.zeros(D)
    hi = rng.uniform(0.5, 2.0, D)
    P = O.Problem(dims[0], dims[1], 3, n_cells, lo, hi, a)
    u = rng.standard_normal(P.n_dofs)
    ours = P.apply(u)
    ref = _apply_collapsed(u, dims, n_cells, hi / np.array(n_cells), a, 3)
    assert np.max(np.abs(ours - ref)) <= 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("k,nq,coll", [(3, 4, False), (3, 5, False),
                                       (3, 4, True), (2, 3, False),
                                       (4, 5, False)])
def test_invariants(k, nq, coll):
    rng = np.random.default_rng(k * 10 + nq)
    a = rng.uniform(-1, 1, 4)
    P = O.Problem(2, 2, k, [3, 2, 2, 3], [0] * 4, [1, 2, 1, 1], a,
                  n_points=nq, collocation=coll)
    # constants are steady
    one = np.ones(P.n_dofs)
    assert np.max(np.abs(P.apply(one))) < 1e-12
    u = rng.standard_normal(P.n_dofs)
    du = P.apply(u)
    # mass: d/dt int u = int du = 0 (periodic); int via the norm machinery
    # is quadratic, so use (||u+du||^2 - ||u-du||^2)/4 = (u, du)_M with u=1
    n_plus = P.norm_and_error(one + du)[0] ** 2
    n_minus = P.norm_and_error(one - du)[0] ** 2
    assert abs(n_plus - n_minus) / 4 < 1e-10
    # upwind dissipation: (u, du)_M <= 0
    ip = (P.norm_and_error(u + du)[0] ** 2 -
          P.norm_and_error(u - du)[0] ** 2) / 4
    assert ip <= 1e-12
    # ECL == FCL
    assert np.max(np.abs(P.apply(u, fcl=True) - du)) <= \
        1e-12 * np.max(np.abs(du))


def test_skew_form_equals_conservative_for_constant_velocity():
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, 4)
    P0 = O.Problem(2, 2, 3, [3, 3, 2, 2], [0] * 4, [1] * 4, a, skew=0.0)
    P1 = O.Problem(2, 2, 3, [3, 3, 2, 2], [0] * 4, [1] * 4, a, skew=0.5)
    u = rng.standard_normal(P0.n_dofs)
    d0, d1 = P0.apply(u), P1.apply(u)
    assert np.max(np.abs(d0 - d1)) <= 1e-12 * np.max(np.abs(d0))


def test_splitmix_known_values():
    # SplitMix64 with state 0: first outputs 0xE220A8397B1DCDAF,
    # 0x6E789E6AA1B965F4 (published reference sequence)
    for i, z in enumerate([0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4]):
        expect = (z >> 11) * (2.0 / 2 ** 53) - 1.0
        assert O.splitmix_uniform(0, i) == expect
