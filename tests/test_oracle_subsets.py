"""Pins of the oracle's disabled-component operators (the V/n/p mask walk
of reference include/templatized_parameters.h:44-101) with a manufactured
solution, after the reference's tests/disabling_components.cpp:148-276.

The reference's test freezes the disabled components at the projection of
the exact solution, solves the enabled ones and checks their L2 errors.  Its
manufactured solution needs a spatially varying linear recombination term;
this one needs none.  On [-1,1]^2 with V = x (E = -1), U_T = 1 and unit
mobilities,
    n = e^x + 1   carries J_n = n E + grad n = -1,
    p = e^-x + 1  carries J_p = -p E - grad p = +1,
both divergence-free, and doping C = n - p makes the Poisson source zero.
Dirichlet data is the exact solution on every side.  Each subset of
{V, n, p} is solved on 4x4 and 8x8 meshes: every enabled component
converges at order k+1 (the error ratio at least 0.7 * 2^(k+1)), the
densities' nodal errors stay within 1.5 times those of the L2 projection of
the exact solution, V = x is recovered to rounding whenever n and p are
frozen, and the disabled components keep their values bitwise.  A mask that updated a disabled
component, dropped a coupling term to a frozen one, or used the wrong
component's rows fails one of these checks; the planted case shows one.
"""
import numpy as np
import pytest

from ddhdg_amd import fe
from ddhdg_amd.problem import (DIRICHLET, KB, Q, Adimensionalizer, BoundaryConditionHandler,
                               FixedTauNPSolverParameters, HomogeneousMobility,
                               HomogeneousPermittivity, HomogeneousProblem,
                               LinearRecombinationTerm, NonlinearSolverParameters, V, n, p)
from tests import helpers as H

EXACT = {V: lambda x: x[..., 0],
         n: lambda x: np.exp(x[..., 0]) + 1.0,
         p: lambda x: np.exp(-x[..., 0]) + 1.0}

SUBSETS = [(True, True, True), (True, False, False), (False, True, False), (False, False, True),
           (True, True, False), (True, False, True), (False, True, True)]


def subset_problem(n_cells, degree, dim=2):
    mesh = fe.BoxMesh(dim, [n_cells] * dim, [-1.0] * dim, [1.0] * dim)
    bh = BoundaryConditionHandler()
    for bid in range(2 * dim):
        for c in (V, n, p):
            bh.add_boundary_condition(bid, DIRICHLET, c, EXACT[c])
    prob = HomogeneousProblem(mesh, HomogeneousPermittivity(Q), HomogeneousMobility(1.0, n),
                              HomogeneousMobility(1.0, p), LinearRecombinationTerm(0, 0, 0),
                              Q / KB, lambda x: EXACT[n](x) - EXACT[p](x), bh, 1.0, 1.0, 0.0, 0.0)
    params = FixedTauNPSolverParameters(degree, degree, degree,
                                        NonlinearSolverParameters(1e-10, 1e-10, 30), Q, 1, 1,
                                        consistent_drift_jacobian=True)
    return prob, params, Adimensionalizer()


def projected_state(mesh, degree, scale=(1.0, 1.0, 1.0)):
    """NPSolver::set_component with use_projection (the reference's
    project_component): cells by L2 projection, trace by the l2 average of
    the cell traces; each component times scale[c]"""
    ref = fe.RefCell(mesh, degree)
    cq = H._cell_qpoints(mesh, ref)
    cfq = fe.cell_face_points(mesh, H._face_qpoints(mesh, ref))
    d = mesh.dim
    cell = np.zeros((mesh.ncells, 3, d + 1, ref.nb))
    tr = np.zeros((mesh.nfaces, 3, ref.nf))
    for c in (V, n, p):
        u, W = fe.project_cells(mesh, ref, scale[c] * EXACT[c](cq), scale[c] * EXACT[c](cfq))
        cell[:, c, d], cell[:, c, :d] = u, W
        tr[:, c] = fe.trace_l2_average(mesh, ref, fe.face_values(mesh, ref, u))
    return cell.ravel(), tr.ravel()


def nodal_errors(cell, mesh, degree):
    from tests.test_oracle_neumann import node_coords
    d = mesh.dim
    c = cell.reshape(mesh.ncells, 3, d + 1, (degree + 1) ** d)
    x = node_coords(mesh, degree)
    return [np.abs(c[:, comp, d] - EXACT[comp](x)).max() for comp in (V, n, p)]


def start_state(mesh, degree, enabled):
    """exact projection for the frozen components, a perturbed one (V + 0.1,
    densities x 0.9) for the solved ones"""
    cell, tr = projected_state(mesh, degree)
    d = mesh.dim
    c = cell.reshape(mesh.ncells, 3, d + 1, -1).copy()
    t = tr.reshape(mesh.nfaces, 3, -1).copy()
    if enabled[V]:
        c[:, V, d] += 0.1
        t[:, V] += 0.1
    for comp in (n, p):
        if enabled[comp]:
            c[:, comp] *= 0.9
            t[:, comp] *= 0.9
    return c.ravel(), t.ravel()


def solve_subset(n_cells, degree, enabled):
    prob, params, adim = subset_problem(n_cells, degree)
    o = H.oracle_for(prob, params, adim, enabled)
    c0, t0 = start_state(prob.mesh, degree, enabled)
    o.set_state(c0, t0)
    res = o.run_newton(enabled=enabled, max_it=30)
    cell, tr = o.get_state()
    return res, (c0, t0), (cell, tr), prob.mesh


@pytest.mark.parametrize("enabled", SUBSETS)
@pytest.mark.parametrize("degree", [1, 2])
def test_subset_converges_at_order_k_plus_1(enabled, degree):
    errs = []
    for nc in (4, 8):
        res, (c0, t0), (cell, tr), mesh = solve_subset(nc, degree, enabled)
        assert res.converged and res.iterations <= 10, (nc, res.iterations)
        d, nb, nf = mesh.dim, (degree + 1) ** mesh.dim, degree + 1
        cc, cs = cell.reshape(mesh.ncells, 3, -1), c0.reshape(mesh.ncells, 3, -1)
        tt, ts = tr.reshape(mesh.nfaces, 3, -1), t0.reshape(mesh.nfaces, 3, -1)
        for comp in (V, n, p):
            if not enabled[comp]:
                # disabled components keep their values bitwise
                assert np.array_equal(cc[:, comp], cs[:, comp])
                assert np.array_equal(tt[:, comp], ts[:, comp])
        e = nodal_errors(cell, mesh, degree)
        # the L2 projection of the exact solution: the frozen components' error
        e_proj = nodal_errors(projected_state(mesh, degree)[0], mesh, degree)
        if enabled[V] and not enabled[n] and not enabled[p]:
            # frozen exact densities: V = x is a discrete solution
            assert e[V] < 1e-10, e
        for comp in (n, p):
            if enabled[comp]:
                assert e[comp] <= 1.5 * e_proj[comp], (comp, e, e_proj)
        errs.append(e)
    for comp in (V, n, p):
        if enabled[comp] and errs[0][comp] > 1e-9:
            ratio = errs[0][comp] / errs[1][comp]
            assert ratio > 0.7 * 2 ** (degree + 1), (comp, errs, ratio)


def test_subset_planted_mask_mistake_fails():
    """the {n} solve run with the mask {V, n} instead (V's rows switched on
    as well) moves the frozen V off its projection: the bitwise check of the
    disabled component fails"""
    prob, params, adim = subset_problem(4, 1)
    c0, t0 = start_state(prob.mesh, 1, (False, True, False))
    o = H.oracle_for(prob, params, adim, (True, True, False))
    o.set_state(c0, t0)
    o.run_newton(enabled=(True, True, False), max_it=30)
    cell, _ = o.get_state()
    cc, cs = cell.reshape(prob.mesh.ncells, 3, -1), c0.reshape(prob.mesh.ncells, 3, -1)
    assert not np.array_equal(cc[:, V], cs[:, V])
