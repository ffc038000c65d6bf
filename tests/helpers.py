"""Shared test helpers: the same problem handed to the oracle and to the
HIP path, and seeded physical states."""
from __future__ import annotations

import numpy as np

from ddhdg_amd import fe
from ddhdg_amd.problem import (DIRICHLET, V, compute_density, evaluate,
                               local_charge_neutrality_potential, n, p)
from oracle import oracle as O


def oracle_for(problem, params, adim, enabled=(True, True, True)):
    """Oracle instance of a ddhdg_amd problem (physical data at its own
    quadrature points)."""
    rec = problem.recombination_term
    from ddhdg_amd.problem import (CellFaceTauNPSolverParameters, LinearRecombinationTerm,
                                   ShockleyReadHallFixedTemperature, flux_type_code)
    cellface = isinstance(params, CellFaceTauNPSolverParameters)
    tau = (0.0, 0.0, 0.0) if cellface else (params.V_tau, params.n_tau, params.p_tau)
    if isinstance(rec, ShockleyReadHallFixedTemperature):
        rtype, rp = 1, (rec.intrinsic_carrier_concentration, rec.electron_life_time,
                        rec.hole_life_time)
    elif isinstance(rec, LinearRecombinationTerm):
        rtype, rp = 2, (rec.constant_term, rec.n_linear_coefficient, rec.p_linear_coefficient)
    else:
        rtype, rp = 0, (0, 0, 0)
    mesh = problem.mesh
    P = O.make_problem(
        mesh.dim, params.V_degree, mesh.n, mesh.lower, mesh.upper,
        epsilon=problem.permittivity.epsilon, mu_n=problem.n_mobility.mu,
        mu_p=problem.p_mobility.mu, temperature=float(problem.temperature),
        nc=problem.conduction_band_density, nv=problem.valence_band_density,
        ec=problem.conduction_band_edge_energy, ev=problem.valence_band_edge_energy,
        recombination=rtype, rec_params=rp, tau=tau,
        flux_type=flux_type_code(params.dd_flux_type), phi_linearize=params.phi_linearize,
        adim=(adim.scale_length, adim.temperature_magnitude, adim.doping_magnitude,
              adim.mobility_magnitude),
        bc_type=problem.boundary_handler.bc_type_table(),
        drift_jacobian=int(params.consistent_drift_jacobian))
    o = O.Oracle(P)
    if cellface:
        o.set_cell_face_tau(params.tau_table(o.ncells, 2 * mesh.dim))
    o.set_enabled(*enabled)
    o.set_doping(evaluate(problem.doping, o.cell_qpoints()))
    fq = o.face_qpoints()
    dirv = np.zeros((o.nfaces, 3, o.Qf))
    neuv = np.zeros((o.nfaces, 3, o.Qf))
    for lf in range(2 * mesh.dim):
        faces = mesh.cell_faces[mesh.cell_bid[:, lf] == lf, lf]
        for c in range(3):
            bc = problem.boundary_handler.get(lf, c)
            if bc is None:
                continue
            kind, f = bc
            (dirv if kind == DIRICHLET else neuv)[faces, c, :] = evaluate(f, fq[faces])
    o.set_bc_values(dirv.ravel(), neuv.ravel())
    return o


def equilibrium_guess(problem, adim, degree):
    """charge-neutral V, Boltzmann n and p (projections on the host)."""
    mesh = problem.mesh
    ref = fe.RefCell(mesh, degree)
    # quadrature points without a GPU: same layout as hdg_cell_qpoints
    o = None
    cq = _cell_qpoints(mesh, ref)
    fq = _face_qpoints(mesh, ref)
    cfq = fq[mesh.cell_faces]
    T = float(problem.temperature)
    d, nb, nf = mesh.dim, ref.nb, ref.nf
    cell = np.zeros((mesh.ncells, 3, d + 1, nb))
    trace = np.zeros((mesh.nfaces, 3, nf))

    def phi(x):
        return local_charge_neutrality_potential(problem, evaluate(problem.doping, x), T)
    Vrf = adim.get_component_rescaling_factor(V)
    u, W = fe.project_cells(mesh, ref, phi(cq) / Vrf, phi(cfq) / Vrf)
    cell[:, V, d], cell[:, V, :d] = u, W
    trace[:, V] = fe.trace_project(mesh, ref, phi(fq) / Vrf)
    for c in (n, p):
        rf = adim.get_component_rescaling_factor(c)
        u, W = fe.project_cells(mesh, ref, compute_density(problem, c, 0.0, phi(cq), T) / rf,
                                compute_density(problem, c, 0.0, phi(cfq), T) / rf)
        cell[:, c, d], cell[:, c, :d] = u, W
        trace[:, c] = fe.trace_project(mesh, ref, compute_density(problem, c, 0.0, phi(fq), T) / rf)
    del o
    return cell.ravel(), trace.ravel()


def perturbed(cell, trace, seed, rel=1e-2):
    rng = np.random.default_rng(seed)
    c = cell * (1 + rel * rng.standard_normal(cell.shape)) + rel * 1e-3 * rng.standard_normal(cell.shape)
    t = trace * (1 + rel * rng.standard_normal(trace.shape))
    return c, t


def _cell_qpoints(mesh, ref):
    d, k = mesh.dim, ref.k
    idx = np.array(np.unravel_index(np.arange(ref.Qc), [k + 1] * d, order="F")).T
    out = np.zeros((mesh.ncells, ref.Qc, d))
    for a in range(d):
        out[:, :, a] = mesh.lower[a] + (mesh.ic[:, a][:, None] + ref.qx[idx[:, a]][None, :]) * mesh.h[a]
    return out


def _face_qpoints(mesh, ref):
    d, k = mesh.dim, ref.k
    out = np.zeros((mesh.nfaces, ref.Qf, d))
    fidx = np.array(np.unravel_index(np.arange(ref.Qf), [k + 1] * (d - 1), order="F")).T
    for a in range(d):
        m = mesh.face_dims[a]
        cnt = mesh.face_off[a + 1] - mesh.face_off[a]
        fi = np.arange(cnt)
        pos = np.stack([fi % m[0], (fi // m[0]) % m[1], fi // (m[0] * m[1])], axis=1)
        for b in range(d):
            out[mesh.face_off[a]:mesh.face_off[a + 1], :, b] = (mesh.lower[b] + pos[:, b] * mesh.h[b])[:, None]
        for mm in range(d - 1):
            b = fe.TANGENTS[d][a][mm]
            out[mesh.face_off[a]:mesh.face_off[a + 1], :, b] += ref.qx[fidx[:, mm]][None, :] * mesh.h[b]
    return out


def rel_fro(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)
