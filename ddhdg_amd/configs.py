"""The silicon diode workloads the path is measured on.

Shared settings (silicon, 300 K): eps_r = 11.7, mu_n = 0.14, mu_p = 0.045
m^2/Vs, Nc = Nv = n_i = 1.45e16 m^-3 with Ec = Ev = 0, SRH with
n_i = 1.45e16 m^-3 and tau_n = tau_p = 1e-6 s, HDG tau = 1 in rescaled units,
use_cell flux, Adimensionalizer(1e-6 m, q/k_B, 1e23 m^-3, 1).  Mesh
coordinates are in micrometres (scale_length units).  Ohmic contacts:
Dirichlet on boundary ids 0 (x = 0) and 1 (x = 1 um), V = charge-neutral
potential (+ bias on the x = 1 contact), n and p the equilibrium densities;
no condition on the other ids (zero normal flux).
"""
from __future__ import annotations

import numpy as np

from .fe import BoxMesh
from .problem import (DIRICHLET, EPSILON0, Adimensionalizer, BoundaryConditionHandler,
                      CellFaceTauNPSolverParameters, FixedTauNPSolverParameters, HomogeneousMobility, HomogeneousPermittivity,
                      HomogeneousProblem, NonlinearSolverParameters,
                      ShockleyReadHallFixedTemperature, compute_density,
                      local_charge_neutrality_potential, n, p, V)

NI = 1.45e16
T_K = 300.0


def step_doping(nd=1e23, na=1e23, junction=0.5):
    def f(x):
        return np.where(x[..., 0] < junction, nd, -na)
    return f


def gaussian_doping(peak=1e24, sigma=0.1, background=1e22, center=(0.5, 0.5, 0.5)):
    c = np.asarray(center)

    def f(x):
        r2 = np.sum((x - c[: x.shape[-1]]) ** 2, axis=-1)
        return background - peak * np.exp(-r2 / (2 * sigma ** 2))
    return f


def scaled_unit_tau(adim):
    """physical tau values whose rescaled values are 1 (the diffusion scale
    D/h of the meshes used here).  The reference default tau = 1 leaves the
    rescaled n, p stabilisation at 1e-12, where the local problems are
    nearly singular (condition ~1e11)."""
    return tuple(adim.get_tau_rescaling_factor(c) for c in (V, n, p))


def diode(dim, n_cells, degree, doping=None, bias=0.0, tau=None, nonlinear=None,
          consistent_drift_jacobian=True, dd_flux_type="use_cell", phi_linearize=False,
          cell_face_tau=None):
    """(problem, parameters, adimensionalizer) of a box diode.

    The diode runs use the consistent linearisation of the volume drift term
    (the reference's own converges at best linearly and diverges on these
    1e23 m^-3 junctions); pass consistent_drift_jacobian=False for the
    reference's.  n_cells: cells per axis (one number for all axes, or one
    per axis).  dd_flux_type and phi_linearize as the reference's
    NPSolverParameters; cell_face_tau: physical tau per (cell, face,
    component), shape (ncells, 2*dim, 3), selects
    CellFaceTauNPSolverParameters."""
    ncs = list(n_cells) if hasattr(n_cells, "__len__") else [n_cells] * dim
    mesh = BoxMesh(dim, ncs, [0.0] * dim, [1.0] * dim)
    doping = doping or step_doping()
    bh = BoundaryConditionHandler()
    prob = HomogeneousProblem(
        mesh=mesh,
        permittivity=HomogeneousPermittivity(11.7 * EPSILON0),
        n_mobility=HomogeneousMobility(0.14, n),
        p_mobility=HomogeneousMobility(0.045, p),
        recombination_term=ShockleyReadHallFixedTemperature(NI, 1e-6, 1e-6),
        temperature=T_K,
        doping=doping,
        boundary_handler=bh,
        conduction_band_density=NI,
        valence_band_density=NI,
    )
    set_contacts(prob, bias)
    adim = Adimensionalizer(1e-6, 1.602176634e-19 / 1.380649e-23, 1e23, 1.0)
    nl = nonlinear or NonlinearSolverParameters(1e-10, 1e-10, 100, 1.0)
    if cell_face_tau is not None:
        tab = np.asarray(cell_face_tau, dtype=np.float64).reshape(mesh.ncells, 2 * dim, 3)
        params = CellFaceTauNPSolverParameters(degree, degree, degree, nl,
                                               dd_flux_type=dd_flux_type,
                                               phi_linearize=phi_linearize)
        m = mesh.ncells * 2 * dim
        params.set_multiple_faces(dim, [0] * m, np.repeat(np.arange(mesh.ncells), 2 * dim),
                                  np.tile(np.arange(2 * dim), mesh.ncells), tab[:, :, 0].ravel(),
                                  tab[:, :, 1].ravel(), tab[:, :, 2].ravel())
    else:
        tau = tau or scaled_unit_tau(adim)
        params = FixedTauNPSolverParameters(degree, degree, degree, nl, tau[0], tau[1], tau[2],
                                            dd_flux_type=dd_flux_type,
                                            phi_linearize=phi_linearize)
    params.consistent_drift_jacobian = consistent_drift_jacobian
    return prob, params, adim


def contact_potential(prob, x):
    return local_charge_neutrality_potential(prob, prob.doping(x), prob.temperature)


def set_contacts(prob, bias=0.0):
    """ohmic contacts on ids 0 and 1; the bias raises the x = 1 contact."""
    bh = prob.boundary_handler

    def Vc(b):
        return lambda x: contact_potential(prob, x) + b

    def dens(c):
        return lambda x: compute_density(prob, c, 0.0, contact_potential(prob, x), prob.temperature)

    for bid, b in ((0, 0.0), (1, bias)):
        bh.add_boundary_condition(bid, DIRICHLET, V, Vc(b))
        bh.add_boundary_condition(bid, DIRICHLET, n, dens(n))
        bh.add_boundary_condition(bid, DIRICHLET, p, dens(p))


def set_bias(solver, bias):
    """bias continuation step: same-type replacement of the x = 1 contact
    potential (NPSolver::replace_boundary_condition)."""
    prob = solver.problem
    solver.replace_boundary_condition(
        1, DIRICHLET, V, lambda x: contact_potential(prob, x) + bias)


def solve_equilibrium_and_bias(solver, biases=(), log=None):
    """The reference's diode workflow (examples/residual.ipynb cells 4-17):
    thermodynamic equilibrium (V only, Boltzmann carriers), densities from
    the potential, n and p with V frozen, then the coupled system; then one
    coupled solve per bias point, the Newton state kept between them.
    Returns the list of NonlinearIterationResults (one per stage)."""
    results = [solver.compute_thermodynamic_equilibrium(log=log)]
    solver.set_enabled_components(False, True, True)
    results.append(solver.run(log=log))
    solver.set_enabled_components(True, True, True)
    results.append(solver.run(log=log))
    for b in biases:
        set_bias(solver, b)
        results.append(solver.run(log=log))
    return results


def batched_devices(count=64, seed=20261203):
    """the batched workload: `count` 2D diodes drawn from `seed`, side
    uniform in 16..192 cells, degree uniform in {1, 2, 3}, donor and acceptor
    densities log-uniform in [1e21, 1e24] m^-3, junction uniform in
    [0.3, 0.7] um, contact bias uniform in [0, 0.6] V.  Returns a list of
    (problem, parameters, adimensionalizer) in draw order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        side = int(rng.integers(16, 193))
        k = int(rng.choice([1, 2, 3]))
        dop = step_doping(10 ** rng.uniform(21, 24), 10 ** rng.uniform(21, 24),
                          rng.uniform(0.3, 0.7))
        out.append(diode(2, side, k, doping=dop, bias=float(rng.uniform(0, 0.6))))
    return out
