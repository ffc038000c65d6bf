"""NPSolver: the Newton driver around the gfx950 hot path.

Mirrors the reference ``NPSolver<dim, Problem>`` host API
(include/np_solver.h:44-914): ``run``, ``compute_thermodynamic_equilibrium``,
``set_current_solution``, ``project_component``, ``set_enabled_components``,
``replace_boundary_condition``, ``assemble_system``.

Per Newton step (``private_run``, src/np_solver.cpp:1452-1561):
  pass 1   hdg_condense      GPU: assembly + condensation + scatter
  solve    global trace system, host sparse direct LU (SciPy SuperLU, the
           stand-in for the reference's SparseDirectUMFPACK; the global
           solve is outside the per-cell hot path)
  pass 2   hdg_reconstruct   GPU: cell update from the trace update
  update   L-infinity norms and the alpha-damped update on the device
The Newton state stays in device memory between steps and between
continuation solves.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import fe
from .hip_lib import HdgContext, MeshDesc, Phys, TraceSolver, gmres_solve, max_abs4
from .problem import (DIRICHLET, EV, KB, NEUMANN, Q, Adimensionalizer,
                      CellFaceTauNPSolverParameters, FixedTauNPSolverParameters,
                      LinearRecombinationTerm, flux_type_code,
                      ShockleyReadHallFixedTemperature, compute_density, evaluate,
                      local_charge_neutrality_potential, n, p, V)


@dataclass
class NonlinearIterationResults:
    """include/nonlinear_iteration_results.h:5-14"""
    converged: bool
    iterations: int
    last_update_norm: float


def fold_physics(problem, parameters, adim) -> Phys:
    """Dimensionless coefficients as the reference ScratchData computes them
    (include/permittivity.h:67-83, include/mobility.h:66-82,
    src/np_solver_parameters.cpp:13-28, include/adimensionalizer.h)."""
    ph = Phys()
    ph.eps = problem.permittivity.epsilon / adim.get_permittivity_rescaling_factor()
    ph.mu_n = problem.n_mobility.mu / adim.get_mobility_rescaling_factor()
    ph.mu_p = problem.p_mobility.mu / adim.get_mobility_rescaling_factor()
    T = float(problem.temperature)
    ph.U_T = (KB * T / Q) / adim.get_thermal_voltage_rescaling_factor()
    if isinstance(parameters, CellFaceTauNPSolverParameters):
        taus = (0.0, 0.0, 0.0)    # per (cell, face): hdg_set_cell_face_tau
    else:
        taus = (parameters.V_tau, parameters.n_tau, parameters.p_tau)
    for c in range(3):
        ph.tau[c] = taus[c] / adim.get_tau_rescaling_factor(c)
    r = problem.recombination_term
    if isinstance(r, ShockleyReadHallFixedTemperature):
        ph.recombination = 1
        ph.rec[0] = r.intrinsic_carrier_concentration
        ph.rec[1] = r.electron_life_time
        ph.rec[2] = r.hole_life_time
    elif isinstance(r, LinearRecombinationTerm):
        ph.recombination = 2
        ph.rec[0], ph.rec[1], ph.rec[2] = (r.constant_term, r.n_linear_coefficient,
                                           r.p_linear_coefficient)
    elif r is None:
        ph.recombination = 0
    else:
        raise NotImplementedError(f"recombination model {type(r).__name__}")
    ph.n_rescale = adim.get_component_rescaling_factor(n)
    ph.rec_rescale = adim.get_recombination_rescaling_factor()
    V_rf = adim.get_component_rescaling_factor(V)
    ph.kT_over_q = KB * T / (Q * V_rf)
    ph.ec_r = problem.conduction_band_edge_energy * EV / (Q * V_rf)
    ph.ev_r = problem.valence_band_edge_energy * EV / (Q * V_rf)
    ph.nc = problem.conduction_band_density
    ph.nv = problem.valence_band_density
    ph.flux_type = flux_type_code(parameters.dd_flux_type)
    ph.phi_linearize = 1 if parameters.phi_linearize else 0
    ph.drift_jacobian = 1 if parameters.consistent_drift_jacobian else 0
    tab = problem.boundary_handler.bc_type_table()
    for b in range(6):
        for c in range(3):
            ph.bc_type[b][c] = tab[b][c]
    return ph


class NPSolver:
    def __init__(self, problem, parameters=None, adimensionalizer=None, device="cuda"):
        import torch
        self.torch = torch
        self.problem = problem
        self.parameters = parameters or FixedTauNPSolverParameters()
        self.adim = adimensionalizer or Adimensionalizer()
        degs = {self.parameters.V_degree, self.parameters.n_degree, self.parameters.p_degree}
        if len(degs) != 1:
            raise NotImplementedError("V, n and p must share one degree")
        self.degree = degs.pop()
        mesh = problem.mesh
        self.mesh = mesh
        self.ref = fe.RefCell(mesh, self.degree)
        md = MeshDesc()
        md.dim, md.degree = mesh.dim, self.degree
        for a in range(3):
            md.n_cells[a] = mesh.n[a]
            md.lower[a] = mesh.lower[a]
            md.upper[a] = mesh.upper[a]
        self.ctx = HdgContext(md, fold_physics(problem, self.parameters, self.adim))
        self.device = torch.device(device)
        self.cell_qp = self.ctx.cell_qpoints()
        self.face_qp = self.ctx.face_qpoints()
        self.enabled = (True, True, True)
        self._upload_problem_data()
        self.x_cell = torch.zeros(self.ctx.ncells * self.ctx.NC, dtype=torch.float64,
                                  device=self.device)
        self.x_trace = torch.zeros(self.ctx.nfaces * 3 * self.ctx.nf, dtype=torch.float64,
                                   device=self.device)
        self._pattern_cache = {}
        self._solver_cache = {}
        self.update_cell = None        # the last Newton update of the cells
        self._tau_version = None
        self._sync_tau()

    # ------------------------------------------------------------ set-up
    def doping_values(self):
        """rescaled doping at the cell quadrature points (what hdg_set_doping
        receives)"""
        return evaluate(self.problem.doping, self.cell_qp) / self.adim.get_component_rescaling_factor(n)

    def _upload_problem_data(self):
        self.ctx.set_doping(self.doping_values())
        self._upload_bc()

    def _upload_bc(self):
        self.ctx.set_bc_values(*self.bc_values())

    def bc_values(self):
        """rescaled Dirichlet and Neumann data at the face quadrature points
        of the boundary faces (what hdg_set_bc_values receives)"""
        ctx, bh = self.ctx, self.problem.boundary_handler
        dirv = np.zeros((ctx.nfaces, 3, ctx.Qf))
        neuv = np.zeros((ctx.nfaces, 3, ctx.Qf))
        mesh = self.mesh
        for lf in range(2 * mesh.dim):
            faces = mesh.cell_faces[mesh.cell_bid[:, lf] == lf, lf]
            for c in range(3):
                bc = bh.get(lf, c)
                if bc is None:
                    continue
                kind, f = bc
                vals = evaluate(f, self.face_qp[faces])
                if kind == DIRICHLET:
                    dirv[faces, c, :] = vals / self.adim.get_component_rescaling_factor(c)
                else:
                    neuv[faces, c, :] = vals / self.adim.get_neumann_boundary_condition_rescaling_factor(c)
        return dirv.ravel(), neuv.ravel()

    def _sync_tau(self):
        """CellFaceTauComputer reads the parameters' tau maps at every
        assembly (include/np_solver_parameters.h:157-236): upload them when
        they changed since the last upload"""
        prm = self.parameters
        if not isinstance(prm, CellFaceTauNPSolverParameters) or prm.version == self._tau_version:
            return
        tab = prm.tau_table(self.ctx.ncells, 2 * self.mesh.dim)
        rf = np.array([self.adim.get_tau_rescaling_factor(c) for c in range(3)])
        self.ctx.set_cell_face_tau(tab / rf)
        self._tau_version = prm.version

    def replace_boundary_condition(self, boundary_id, bc_type, component, f):
        """NPSolver::replace_boundary_condition (src/np_solver.cpp:2160-2193)."""
        old = self.problem.boundary_handler.get(boundary_id, component)
        if old is None or old[0] != bc_type:
            raise NotImplementedError("only same-type replacement is supported")
        self.problem.boundary_handler.replace_boundary_condition(boundary_id, bc_type,
                                                                 component, f)
        self._upload_bc()

    def set_enabled_components(self, V_enabled, n_enabled, p_enabled):
        self.enabled = (bool(V_enabled), bool(n_enabled), bool(p_enabled))
        self.ctx.set_enabled(*self.enabled)

    def is_enabled(self, c):
        return self.enabled[c]

    # ----------------------------------------------------------- state
    @property
    def NC(self):
        return self.ctx.NC

    def get_state(self):
        return (self.x_cell.cpu().numpy().copy(), self.x_trace.cpu().numpy().copy())

    def set_state(self, cell, trace):
        t = self.torch
        self.x_cell.copy_(t.as_tensor(np.asarray(cell, dtype=np.float64).ravel()))
        self.x_trace.copy_(t.as_tensor(np.asarray(trace, dtype=np.float64).ravel()))

    def _write_component(self, c, u, W, trace):
        cell, tr = self.get_state()
        d, nb, nf = self.mesh.dim, self.ref.nb, self.ref.nf
        cell = cell.reshape(self.ctx.ncells, 3, d + 1, nb)
        cell[:, c, d, :] = u
        cell[:, c, :d, :] = W
        tr = tr.reshape(self.ctx.nfaces, 3, nf)
        tr[:, c, :] = trace
        self.set_state(cell, tr)

    def project_component(self, c, f):
        """NPSolver::project_component for V, n or p (src/project_component.cpp:
        478-574): cells by L2 projection, trace by l2_average."""
        rf = self.adim.get_component_rescaling_factor(c)
        fc = evaluate(f, self.cell_qp) / rf
        ff = evaluate(f, fe.cell_face_points(self.mesh, self.face_qp)) / rf
        u, W = fe.project_cells(self.mesh, self.ref, fc, ff)
        tr = fe.trace_l2_average(self.mesh, self.ref, fe.face_values(self.mesh, self.ref, u))
        self._write_component(c, u, W, tr)

    set_component = project_component

    def set_current_solution(self, V_function, n_function, p_function, use_projection=True):
        for c, f in ((V, V_function), (n, n_function), (p, p_function)):
            self.project_component(c, f)

    def _V_at(self, u_V):
        rf = self.adim.get_component_rescaling_factor(V)
        return (np.einsum("qj,cj->cq", self.ref.phic, u_V) * rf,
                fe.face_values(self.mesh, self.ref, u_V) * rf)

    def project_densities_from_potential(self, phi_n=0.0, phi_p=0.0):
        """project_component(phi_n / phi_p, f) (src/project_component.cpp:
        282-452): n, p from the current V and given quasi-Fermi potentials."""
        cell, _ = self.get_state()
        d, nb = self.mesh.dim, self.ref.nb
        u_V = cell.reshape(self.ctx.ncells, 3, d + 1, nb)[:, V, d, :]
        Vc, Vf = self._V_at(u_V)
        T = float(self.problem.temperature)
        for c, phi in ((n, phi_n), (p, phi_p)):
            rf = self.adim.get_component_rescaling_factor(c)
            pc = evaluate(phi, self.cell_qp)
            pf = evaluate(phi, fe.cell_face_points(self.mesh, self.face_qp))
            fc = compute_density(self.problem, c, pc, Vc, T) / rf
            ff = compute_density(self.problem, c, pf, Vf, T) / rf
            u, W = fe.project_cells(self.mesh, self.ref, fc, ff)
            tr = fe.trace_l2_average(self.mesh, self.ref, fe.face_values(self.mesh, self.ref, u))
            self._write_component(c, u, W, tr)

    def compute_local_charge_neutrality(self):
        """src/thermodynamic_equilibrium.cpp:47-351, 351-359."""
        T = float(self.problem.temperature)
        rf = self.adim.get_component_rescaling_factor(V)

        def phi(x):
            return local_charge_neutrality_potential(
                self.problem, evaluate(self.problem.doping, x), T) / rf

        fc = phi(self.cell_qp)
        ff = phi(fe.cell_face_points(self.mesh, self.face_qp))
        u, W = fe.project_cells(self.mesh, self.ref, fc, ff)
        tr = fe.trace_project(self.mesh, self.ref, phi(self.face_qp))
        self._write_component(V, u, W, tr)

    # ---------------------------------------------------------- Newton
    def _pattern(self):
        key = self.enabled
        if key not in self._pattern_cache:
            rp, ci = self.ctx.csr_pattern()
            nf = self.ctx.nf
            comps = [c for c in range(3) if self.enabled[c]]
            ne = len(comps)
            f = np.repeat(np.arange(self.ctx.nfaces), ne * nf)
            cidx = np.tile(np.repeat(np.asarray(comps), nf), self.ctx.nfaces)
            j = np.tile(np.arange(nf), self.ctx.nfaces * ne)
            r2f = f * 3 * nf + cidx * nf + j
            self._pattern_cache[key] = (rp, ci, self.ctx.dirichlet_dofs(),
                                        self.torch.as_tensor(r2f, device=self.device))
        return self._pattern_cache[key]

    def assemble_system(self, thermodyn=False):
        """pass 1 only; returns the condensed trace matrix (SciPy CSR), the
        right-hand side (= minus the residual) and the Dirichlet values, as
        pyddhdg's assemble_system/get_residual/get_jacobian expose them."""
        import scipy.sparse as sp
        t = self.torch
        self._sync_tau()
        rp, ci, dirs, _ = self._pattern()
        vals = t.empty(self.ctx.nnz, dtype=t.float64, device=self.device)
        rhs = t.empty(self.ctx.ntrace, dtype=t.float64, device=self.device)
        dirv = t.zeros(self.ctx.ntrace, dtype=t.float64, device=self.device)
        self.ctx.condense(self.x_cell, self.x_trace, thermodyn, vals, rhs, dirv)
        A = sp.csr_matrix((vals.cpu().numpy(), ci, rp), shape=(self.ctx.ntrace,) * 2)
        return A, rhs.cpu().numpy(), dirs, dirv.cpu().numpy()[dirs]

    def trace_solver(self):
        """the GPU trace solver for the current enabled components (built on
        first use, kept per component set)."""
        key = self.enabled
        if key not in self._solver_cache:
            self._solver_cache[key] = TraceSolver(self.ctx)
        return self._solver_cache[key]

    def release_trace_solvers(self, keep_current=True):
        """free the device memory (the factors) of the cached trace solvers:
        all of them, or all but the current component set's"""
        for key in [k for k in self._solver_cache if not (keep_current and k == self.enabled)]:
            self._solver_cache.pop(key).close()

    def _device_pattern(self):
        """the trace CSR pattern of the enabled components on the device
        (row pointers int64, column indices int32), checked on the host"""
        key = ("dev",) + tuple(self.enabled)
        if key not in self._pattern_cache:
            rp, ci, _, _ = self._pattern()
            n_ = self.ctx.ntrace
            if (rp[0] != 0 or rp[-1] != len(ci) or np.any(np.diff(rp) < 0) or
                    (len(ci) and (ci.min() < 0 or ci.max() >= n_))):
                raise RuntimeError("trace CSR pattern out of range")
            t = self.torch
            self._pattern_cache[key] = (t.as_tensor(rp, device=self.device),
                                        t.as_tensor(ci, device=self.device))
        return self._pattern_cache[key]

    def _solve_trace(self, vals, rhs, sol):
        """solve_linear_problem (src/np_solver.cpp:1326-1354): the condensed
        trace system, by GMRES on the device when iterative_linear_solver is
        set (restarted every 28 iterations, no preconditioner, tolerance
        1e-10 |rhs|_2, at most 10 n iterations, as the reference's
        SolverGMRES with its default 30 temporary vectors), else by the direct
        solver, on the device or with SuperLU on the host."""
        if self.parameters.iterative_linear_solver:
            rp, ci = self._device_pattern()
            tol = 1e-10 * float(self.torch.linalg.vector_norm(rhs))
            self.gmres_iterations, _ = gmres_solve(rp, ci, vals, rhs, sol, 28, tol)
            return
        if self.parameters.trace_solver == "gpu":
            ts = self.trace_solver()
            ts.factor(vals)
            ts.solve(rhs, sol)
            return
        if self.parameters.trace_solver != "host":
            raise ValueError(f"trace_solver must be 'gpu' or 'host', not "
                             f"{self.parameters.trace_solver!r}")
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        rp, ci, _, _ = self._pattern()
        n_ = self.ctx.ntrace
        A = sp.csc_matrix(sp.csr_matrix((vals.cpu().numpy(), ci, rp), shape=(n_, n_)))
        sol.copy_(self.torch.as_tensor(spla.splu(A).solve(rhs.cpu().numpy())))

    def _private_run(self, abs_tol, rel_tol, max_it, thermodyn, log=None):
        """NPSolver::private_run (src/np_solver.cpp:1483-1587): Newton with the
        L-infinity update criterion against the state before the update.
        Every vector stays on the device; one host synchronisation per
        iteration reads the norms."""
        it = NewtonIteration(self, thermodyn, abs_tol, rel_tol, max_it, log)
        while True:
            it.launch()
            it.sync()
            if it.finish() != "continue":
                break
        return it.result()

    def get_local_cell_system(self, cell_level, cell_index, compute_thermodynamic_equilibrium=False):
        """NPSolver::get_local_cell_system (src/np_solver.cpp:2337-2445): the
        cell's local blocks as the per-cell operator assembles them before
        its Schur complement (after the LocalCondenser), computed on the
        device: ({"cc", "ct", "tc", "tt"}, {"c", "t"}) as numpy arrays (tc
        negated, as the reference stores it).  The mesh has one level; a
        thermodynamic-equilibrium call runs with V alone, as the reference's
        does."""
        if cell_level != 0:
            raise ValueError("the structured mesh has one level (cell_level 0)")
        t = self.torch
        saved = self.enabled
        if compute_thermodynamic_equilibrium:
            self.set_enabled_components(True, False, False)
        try:
            ctx = self.ctx
            N, T = ctx.ne * (self.mesh.dim + 1) * ctx.nb, ctx.T
            dev = self.device
            cells = t.tensor([int(cell_index)], dtype=t.int32, device=dev)
            out = [t.empty(sz, dtype=t.float64, device=dev)
                   for sz in (N * N, N * T, T * N, T * T, N, T)]
            ctx.local_system(self.x_cell, self.x_trace, compute_thermodynamic_equilibrium, cells,
                             *out)
            A, CT, TC, TT, c, tv = (x.cpu().numpy() for x in out)
        finally:
            if compute_thermodynamic_equilibrium:
                self.set_enabled_components(*saved)
        return ({"cc": A.reshape(N, N), "ct": CT.reshape(N, T), "tc": TC.reshape(T, N),
                 "tt": TT.reshape(T, T)}, {"c": c, "t": tv})

    def output_results(self, solution_filename, trace_filename=None, save_update=False,
                       redimensionalize_quantities=True):
        """NPSolver::output_results (src/np_solver.cpp:1737-2033): the cell
        fields (and the last Newton update with save_update) as a VTK .vtu
        file; with trace_filename also the trace dofs."""
        from . import output
        cell, tr = self.get_state()
        upd = None
        if save_update:
            if self.update_cell is None:
                raise RuntimeError("output_results: no Newton update to save yet")
            upd = self.update_cell.cpu().numpy()
        output.write_cells(solution_filename, self.mesh, self.degree, cell, self.adim, upd,
                           redimensionalize_quantities, self.problem.doping)
        if trace_filename is not None:
            output.write_trace(trace_filename, self.mesh, self.degree, tr, self.adim,
                               redimensionalize_quantities)

    def current_scale(self):
        """dimensionless flux integral -> physical current: q mu_mag N V_rf
        l^(dim-2) with V_rf = k_B T_mag / q (A per metre of depth in 2D)."""
        a = self.adim
        return (a.mobility_magnitude * a.doping_magnitude * KB * a.temperature_magnitude *
                a.scale_length ** (self.mesh.dim - 2))

    def terminal_current(self, boundary_id):
        """Current leaving the device through the faces of boundary_id, from
        the HDG numerical fluxes of n and p on those faces (the reference
        has no contact-current integral: a quantity this build defines)."""
        t = self.torch
        out = t.empty(2, dtype=t.float64, device=self.device)
        self.ctx.terminal_current(self.x_cell, self.x_trace, boundary_id, out)
        fn, fp = out.cpu().numpy()
        return float(fn - fp) * self.current_scale()

    def run(self, absolute_tol=None, relative_tol=None, max_number_of_iterations=None, log=None):
        nl = self.parameters.nonlinear_parameters
        return self._private_run(
            nl.absolute_tolerance if absolute_tol is None else absolute_tol,
            nl.relative_tolerance if relative_tol is None else relative_tol,
            nl.max_number_of_iterations if max_number_of_iterations is None
            else max_number_of_iterations, False, log)

    def compute_thermodynamic_equilibrium(self, generate_first_guess=True, absolute_tol=None,
                                          relative_tol=None, max_number_of_iterations=None,
                                          log=None):
        """src/thermodynamic_equilibrium.cpp:365-405."""
        nl = self.parameters.nonlinear_parameters
        saved = self.enabled
        self.set_enabled_components(True, False, False)
        if generate_first_guess:
            self.compute_local_charge_neutrality()
        res = self._private_run(
            nl.absolute_tolerance if absolute_tol is None else absolute_tol,
            nl.relative_tolerance if relative_tol is None else relative_tol,
            nl.max_number_of_iterations if max_number_of_iterations is None
            else max_number_of_iterations, True, log)
        self.project_densities_from_potential(0.0, 0.0)
        self.set_enabled_components(*saved)
        return res


class NewtonIteration:
    """One Newton solve of an NPSolver split into device-side iterations, so
    that a host loop can interleave several solvers (one stream each):
    launch() enqueues a whole iteration on the current stream (pass 1, trace
    factorisation and solve, Dirichlet overwrite, pass 2, update, norms into
    pinned host memory); after the stream has been synchronised, finish()
    applies the reference's L-infinity criterion (src/np_solver.cpp:
    1521-1560) and returns "converged", "continue" or "failed" (iteration
    limit)."""

    def __init__(self, solver, thermodyn, abs_tol, rel_tol, max_it, log=None, graph=None):
        t = solver.torch
        # hipGraph: the first iteration runs eagerly (it also creates the
        # trace solver and the BLAS workspaces), the second is captured and
        # every later one replays it; all buffers below are fixed for the
        # object's lifetime and boundary data are updated in place
        self.use_graph = solver.parameters.graph_newton if graph is None else graph
        if self.use_graph and solver.parameters.iterative_linear_solver:
            raise ValueError("graph_newton needs the direct trace solver: GMRES "
                             "synchronises with the host every iteration")
        self.graph = None
        # capture needs a stream other than the default one
        cur = t.cuda.current_stream()
        self.stream = (t.cuda.Stream() if self.use_graph and cur == t.cuda.default_stream()
                       else cur)
        self.s, self.t = solver, t
        self.thermodyn, self.abs_tol, self.rel_tol = thermodyn, abs_tol, rel_tol
        self.max_it = max_it if max_it >= 0 else 1 << 30
        self.log = log
        ctx, dev = solver.ctx, solver.device
        solver._sync_tau()
        _, _, dirs, self.r2f = solver._pattern()
        self.vals = t.empty(ctx.nnz, dtype=t.float64, device=dev)
        self.rhs = t.empty(ctx.ntrace, dtype=t.float64, device=dev)
        self.sol = t.empty(ctx.ntrace, dtype=t.float64, device=dev)
        self.dirv = t.zeros(ctx.ntrace, dtype=t.float64, device=dev)
        self.upd_trace = t.zeros_like(solver.x_trace)
        self.upd_cell = t.zeros_like(solver.x_cell)
        self.dirs = t.as_tensor(dirs, device=dev)
        self.norms = t.zeros(4, dtype=t.float64).pin_memory()
        self.dnorms = t.zeros(4, dtype=t.float64, device=dev)
        self.nscratch = t.empty(1024, dtype=t.float64, device=dev)
        self.alpha = solver.parameters.nonlinear_parameters.alpha
        self.step, self.converged, self.upd_cell_norm = 0, False, 0.0

    def reset(self):
        """a new Newton solve with the same buffers (and captured graph): the
        boundary data may have changed in place, the components and the
        operator not (per-face tau values are updated in place)"""
        self.s._sync_tau()
        self.step, self.converged, self.upd_cell_norm = 0, False, 0.0

    def launch(self):
        t = self.t
        if self.stream != t.cuda.current_stream():
            self.stream.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(self.stream):
            if not self.use_graph or (self.step == 0 and self.graph is None):
                self._iteration()
            else:
                if self.graph is None:
                    self.graph = t.cuda.CUDAGraph()
                    with t.cuda.graph(self.graph, stream=self.stream):
                        self._iteration()
                self.graph.replay()
        self.step += 1

    def sync(self):
        self.stream.synchronize()

    def _iteration(self):
        s, t = self.s, self.t
        s.ctx.condense(s.x_cell, s.x_trace, self.thermodyn, self.vals, self.rhs, self.dirv)
        s._solve_trace(self.vals, self.rhs, self.sol)
        self.sol[self.dirs] = self.dirv[self.dirs]
        self.upd_trace.zero_()
        self.upd_trace[self.r2f] = self.sol
        s.ctx.reconstruct(s.x_cell, s.x_trace, self.upd_trace, self.thermodyn, self.upd_cell)
        s.update_cell = self.upd_cell
        max_abs4([self.upd_cell, self.upd_trace, s.x_cell, s.x_trace], self.nscratch,
                 self.dnorms)
        self.norms.copy_(self.dnorms, non_blocking=True)
        s.x_cell.add_(self.upd_cell, alpha=self.alpha)
        s.x_trace.add_(self.upd_trace, alpha=self.alpha)

    def finish(self):
        n_ = self.norms.numpy()
        self.upd_cell_norm = float(n_[0])
        upd_norm = max(self.upd_cell_norm, float(n_[1]))
        cur_norm = max(float(n_[2]), float(n_[3]))
        if (not np.isfinite(upd_norm) and self.s.parameters.trace_solver == "gpu"
                and not self.s.parameters.iterative_linear_solver):
            self.s.trace_solver().factor(self.vals, check=True)
        if self.log is not None:
            self.log(self.step, upd_norm, cur_norm)
        if upd_norm < self.abs_tol or upd_norm < self.rel_tol * cur_norm:
            self.converged = True
            return "converged"
        if not np.isfinite(upd_norm) or self.step >= self.max_it:
            return "failed"
        return "continue"

    def result(self):
        return NonlinearIterationResults(self.converged, self.step, self.upd_cell_norm)
