"""Pins of the oracle's Neumann data path (apply_nbc_on_face,
reference include/assembling_methods.template.h:1695-1734, and its call in
assemble_flux_conditions, :2164-2240) against exact discrete solutions.

The reference's own test (tests/neumann_boundary_conditions.cpp:80-187)
puts V Dirichlet on boundary id 0 of [-1,1]^d, V Neumann data on id 1 and
zero V Neumann data on the other ids, n = p = 0 Dirichlet, degree 2, and
expects V = x to L2 < 1e-10 in at most three Newton iterations.  The tests
below keep that structure and strengthen it:
  * the potential is quadratic (V = x^2 + xy (+ yz)) at degree >= 2, so
    the Neumann data varies along every Neumann face, and linear at
    degree 1; every face but id 0 carries nonzero data;
  * the permittivity is q F/m (rescaled permittivity 1) so that every
    block of the local systems is O(1); the reference's literal data
    (1 F/m, tau 1) is checked separately with a row-equilibrated solve;
  * n and p Neumann data are pinned with a drift current that leaves
    through a Neumann contact;
  * in thermodynamic equilibrium the reference replaces the Neumann data by
    the zero function (:2229-2236).
A wrong sign, a missing rescaling factor or data on the wrong face turns
each exact solution into a wrong one; the planted cases show that.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from ddhdg_amd import fe
from ddhdg_amd.problem import (DIRICHLET, KB, NEUMANN, Q, Adimensionalizer,
                               BoundaryConditionHandler, FixedTauNPSolverParameters,
                               HomogeneousMobility, HomogeneousPermittivity,
                               HomogeneousProblem, LinearRecombinationTerm,
                               NonlinearSolverParameters, V, n, p)
from tests import helpers as H


def potential(dim, degree):
    """exact potential and its gradient: quadratic at degree >= 2 (Laplacian 2),
    linear at degree 1"""
    if degree >= 2:
        def Vf(x):
            v = x[..., 0] ** 2 + x[..., 0] * x[..., 1]
            return v + (x[..., 1] * x[..., 2] if dim == 3 else 0.0)

        def grad(x):
            g = np.zeros(x.shape)
            g[..., 0] = 2 * x[..., 0] + x[..., 1]
            g[..., 1] = x[..., 0] + (x[..., 2] if dim == 3 else 0.0)
            if dim == 3:
                g[..., 2] = x[..., 1]
            return g
        return Vf, grad, 2.0
    a = np.array([1.0, 0.5, 0.25][:dim])
    return (lambda x: x[..., :dim] @ a), (lambda x: np.broadcast_to(a, x.shape).copy()), 0.0


def neumann_potential_problem(dim, n_cells, degree, eps=Q, tauV=Q, data_scale=1.0):
    """V Dirichlet on id 0, V Neumann h = eps E.nu = -eps grad V . nu on ids
    1..2d-1 (times data_scale, 1 for the exact data), n = p = 0 Dirichlet;
    doping -eps Lap V / q so that -div(eps grad V) = q C holds."""
    Vf, grad, lap = potential(dim, degree)
    mesh = fe.BoxMesh(dim, [n_cells] * dim, [-1.0] * dim, [1.0] * dim)
    bh = BoundaryConditionHandler()
    for bid in range(2 * dim):
        bh.add_boundary_condition(bid, DIRICHLET, n, 0.0)
        bh.add_boundary_condition(bid, DIRICHLET, p, 0.0)
    bh.add_boundary_condition(0, DIRICHLET, V, Vf)
    for bid in range(1, 2 * dim):
        a, s = bid // 2, (1.0 if bid % 2 else -1.0)
        bh.add_boundary_condition(
            bid, NEUMANN, V, lambda x, a=a, s=s: -data_scale * eps * s * grad(x)[..., a])
    doping = -eps * lap / Q
    prob = HomogeneousProblem(mesh, HomogeneousPermittivity(eps), HomogeneousMobility(1.0, n),
                              HomogeneousMobility(1.0, p), LinearRecombinationTerm(0, 0, 0),
                              Q / KB, doping, bh, 1.0, 1.0, 0.0, 0.0)
    params = FixedTauNPSolverParameters(degree, degree, degree,
                                        NonlinearSolverParameters(1e-10, 1e-10, 10), tauV, 1, 1)
    return prob, params, Adimensionalizer()


def node_coords(mesh, degree):
    """coordinates of the equidistant scalar nodes of every cell"""
    d = mesh.dim
    nb = (degree + 1) ** d
    idx = np.array(np.unravel_index(np.arange(nb), [degree + 1] * d, order="F")).T
    x = np.zeros((mesh.ncells, nb, d))
    for a in range(d):
        x[..., a] = mesh.lower[a] + (mesh.ic[:, a][:, None] + idx[None, :, a] / degree) * mesh.h[a]
    return x


def potential_errors(cell, mesh, degree):
    Vf, grad, _ = potential(mesh.dim, degree)
    d = mesh.dim
    nb = (degree + 1) ** d
    c = cell.reshape(mesh.ncells, 3, d + 1, nb)
    x = node_coords(mesh, degree)
    eV = np.abs(c[:, V, d, :] - Vf(x)).max()
    eE = np.abs(c[:, V, :d, :].transpose(0, 2, 1) + grad(x)).max()   # E = -grad V
    enp = max(np.abs(c[:, n]).max(), np.abs(c[:, p]).max())
    return eV, eE, enp


NEU_CASES = [(2, 2, 1), (2, 2, 2), (2, 4, 2), (2, 3, 3), (3, 1, 1), (3, 1, 2), (3, 2, 2)]


@pytest.mark.parametrize("dim,n_cells,degree", NEU_CASES)
def test_neumann_potential_exact(dim, n_cells, degree):
    """reference tests/neumann_boundary_conditions.cpp:160-187: the exact
    potential in at most three iterations, to 1e-10"""
    prob, params, adim = neumann_potential_problem(dim, n_cells, degree)
    o = H.oracle_for(prob, params, adim)
    res = o.run_newton(max_it=10)
    assert res.converged and res.iterations <= 3
    eV, eE, enp = potential_errors(o.get_state()[0], prob.mesh, degree)
    assert eV < 1e-10 and eE < 1e-10 and enp < 1e-10, (eV, eE, enp)


@pytest.mark.parametrize("scale", [-1.0, 1.0 / Q, 0.5])
def test_neumann_potential_planted_mistakes_fail(scale):
    """the same problem with the Neumann data negated, left unrescaled or
    halved: the solution is no longer the exact potential"""
    prob, params, adim = neumann_potential_problem(2, 2, 2, data_scale=scale)
    o = H.oracle_for(prob, params, adim)
    o.run_newton(max_it=10)
    eV, eE, _ = potential_errors(o.get_state()[0], prob.mesh, 2)
    assert eV > 1e-3 and eE > 1e-3


@pytest.mark.parametrize("dim,n_cells,degree", [(2, 2, 2), (3, 1, 2)])
def test_neumann_reference_literal_data(dim, n_cells, degree):
    """The reference's literal data: permittivity 1 F/m, tau_V = 1, V Neumann
    data -1 on id 1 and 0 on the others, V = x.  Rescaled, the potential rows
    are 6e18 times the flux rows, which unscaled partial pivoting (the
    reference's gauss_jordan) cannot eliminate in FP64; a row-equilibrated
    dense solve of the oracle's uncondensed system recovers V = x."""
    from tests.test_oracle import _full_system
    mesh = fe.BoxMesh(dim, [n_cells] * dim, [-1.0] * dim, [1.0] * dim)
    bh = BoundaryConditionHandler()
    for bid in range(2 * dim):
        bh.add_boundary_condition(bid, DIRICHLET, n, 0.0)
        bh.add_boundary_condition(bid, DIRICHLET, p, 0.0)
    bh.add_boundary_condition(0, DIRICHLET, V, -1.0)
    bh.add_boundary_condition(1, NEUMANN, V, -1.0)
    for bid in range(2, 2 * dim):
        bh.add_boundary_condition(bid, NEUMANN, V, 0.0)
    prob = HomogeneousProblem(mesh, HomogeneousPermittivity(1.0), HomogeneousMobility(1.0, n),
                              HomogeneousMobility(1.0, p), LinearRecombinationTerm(0, 0, 0),
                              Q / KB, 0.0, bh, 1.0, 1.0, 0.0, 0.0)
    params = FixedTauNPSolverParameters(degree, degree, degree,
                                        NonlinearSolverParameters(1e-10, 1e-10, 3))
    o = H.oracle_for(prob, params, Adimensionalizer())
    M, rhs, _ = _full_system(o)
    Md = M.toarray()
    D = 1.0 / np.abs(Md).max(axis=1)
    assert D.max() / D.min() > 1e18
    x = np.linalg.solve(Md * D[:, None], rhs * D)
    d = dim
    nb = (degree + 1) ** d
    c = x[: o.ncells * o.N].reshape(o.ncells, 3, d + 1, nb)
    xn = node_coords(mesh, degree)
    assert np.abs(c[:, V, d] - xn[..., 0]).max() < 1e-10
    assert np.abs(c[:, V, 0] + 1.0).max() < 1e-10
    assert np.abs(c[:, n]).max() < 1e-10 and np.abs(c[:, p]).max() < 1e-10


N0, P0 = 1.0, 2.0


def neumann_carrier_problem(dim, n_cells, degree, sn=1.0, sp_=1.0):
    """V = x between Dirichlet contacts on ids 0 and 1; n = N0 and p = P0
    constant, carried by a drift current.  n and p are Dirichlet on id 0 and
    Neumann on id 1 with the exact numerical fluxes F_n = n mu_n E.nu,
    F_p = -p mu_p E.nu times the Neumann rescaling factor (sn, sp_ = 1);
    zero Neumann data on the lateral ids, where E.nu = 0.  Doping N0 - P0
    keeps the Poisson source zero."""
    mesh = fe.BoxMesh(dim, [n_cells] * dim, [-1.0] * dim, [1.0] * dim)
    adim = Adimensionalizer()
    rf = adim.get_neumann_boundary_condition_rescaling_factor(n)
    bh = BoundaryConditionHandler()
    for bid in (0, 1):
        bh.add_boundary_condition(bid, DIRICHLET, V, lambda x: x[..., 0])
    bh.add_boundary_condition(0, DIRICHLET, n, N0)
    bh.add_boundary_condition(0, DIRICHLET, p, P0)
    e_nu = -1.0                      # E = -grad V = (-1, 0, 0), nu = +x on id 1
    bh.add_boundary_condition(1, NEUMANN, n, sn * rf * N0 * e_nu)
    bh.add_boundary_condition(1, NEUMANN, p, sp_ * rf * (-P0) * e_nu)
    for bid in range(2, 2 * dim):
        for c in (V, n, p):
            bh.add_boundary_condition(bid, NEUMANN, c, 0.0)
    prob = HomogeneousProblem(mesh, HomogeneousPermittivity(Q), HomogeneousMobility(1.0, n),
                              HomogeneousMobility(1.0, p), LinearRecombinationTerm(0, 0, 0),
                              Q / KB, N0 - P0, bh, 1.0, 1.0, 0.0, 0.0)
    params = FixedTauNPSolverParameters(degree, degree, degree,
                                        NonlinearSolverParameters(1e-10, 1e-10, 30), Q, 1, 1,
                                        consistent_drift_jacobian=True)
    return prob, params, adim


def carrier_start(mesh, degree):
    d, nb, nf = mesh.dim, (degree + 1) ** mesh.dim, (degree + 1) ** (mesh.dim - 1)
    cell = np.zeros((mesh.ncells, 3, d + 1, nb))
    tr = np.zeros((mesh.nfaces, 3, nf))
    cell[:, n, d], cell[:, p, d] = 0.8, 2.3
    tr[:, n], tr[:, p] = 0.8, 2.3
    return cell.ravel(), tr.ravel()


def carrier_errors(cell, mesh, degree):
    d = mesh.dim
    c = cell.reshape(mesh.ncells, 3, d + 1, (degree + 1) ** d)
    return (np.abs(c[:, n, d] - N0).max(), np.abs(c[:, p, d] - P0).max(),
            np.abs(c[:, V, 0] + 1.0).max())


@pytest.mark.parametrize("dim,n_cells,degree", [(2, 2, 1), (2, 3, 2), (2, 2, 3), (3, 2, 1), (3, 1, 2)])
def test_neumann_carriers_exact(dim, n_cells, degree):
    prob, params, adim = neumann_carrier_problem(dim, n_cells, degree)
    o = H.oracle_for(prob, params, adim)
    o.set_state(*carrier_start(prob.mesh, degree))
    res = o.run_newton(max_it=30)
    assert res.converged
    en, ep, eE = carrier_errors(o.get_state()[0], prob.mesh, degree)
    assert en < 1e-10 and ep < 1e-10 and eE < 1e-10, (en, ep, eE)


@pytest.mark.parametrize("sn,sp_", [(-1.0, 1.0), (1.0, -1.0), (1.0 / Q, 1.0)])
def test_neumann_carriers_planted_mistakes_fail(sn, sp_):
    prob, params, adim = neumann_carrier_problem(2, 2, 1, sn, sp_)
    o = H.oracle_for(prob, params, adim)
    o.set_state(*carrier_start(prob.mesh, 1))
    o.run_newton(max_it=30)
    en, ep, _ = carrier_errors(o.get_state()[0], prob.mesh, 1)
    assert max(en, ep) > 1e-3


@pytest.mark.parametrize("dim,n_cells,degree", [(2, 2, 2), (3, 1, 1)])
def test_equilibrium_ignores_neumann_data(dim, n_cells, degree):
    """template.h:2229-2236: the equilibrium operator applies the zero
    function on Neumann faces, so the data does not enter; the full operator
    adds it to the trace rhs of the Neumann faces only"""
    prob, params, adim = neumann_potential_problem(dim, n_cells, degree)
    prob0, _, _ = neumann_potential_problem(dim, n_cells, degree, data_scale=0.0)
    rng = np.random.default_rng(1)
    o = H.oracle_for(prob, params, adim, (True, False, False))
    o0 = H.oracle_for(prob0, params, adim, (True, False, False))
    cell = rng.standard_normal(o.ncells * o.NC)
    tr = rng.standard_normal(o.nfaces * 3 * o.nf)
    for oo in (o, o0):
        oo.set_state(cell, tr)
    for c in range(o.ncells):
        S, g, *_ = o.condensed_cell(c, True)
        S0, g0, *_ = o0.condensed_cell(c, True)
        assert np.array_equal(S, S0) and np.array_equal(g, g0)
    o = H.oracle_for(prob, params, adim)
    o0 = H.oracle_for(prob0, params, adim)
    for oo in (o, o0):
        oo.set_state(cell, tr)
    changed = 0
    for c in range(o.ncells):
        S, g, *_ = o.condensed_cell(c, False)
        S0, g0, *_ = o0.condensed_cell(c, False)
        assert np.array_equal(S, S0)
        changed += not np.array_equal(g, g0)
    assert changed > 0
