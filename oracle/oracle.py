"""Python side of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (the C++ restatement of the reference per-cell
path, ``hdg_oracle.cpp``) and restates the reference Newton driver
``NPSolver::private_run`` (src/np_solver.cpp:1452-1561) on top of it, with
SciPy's SuperLU standing in for ``SparseDirectUMFPACK``
(src/np_solver.cpp:1347-1352).  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcProblem(ct.Structure):
    _fields_ = [
        ("dim", ct.c_int),
        ("degree", ct.c_int),
        ("n_cells", ct.c_int * 3),
        ("lower", ct.c_double * 3),
        ("upper", ct.c_double * 3),
        ("epsilon", ct.c_double),
        ("mu_n", ct.c_double),
        ("mu_p", ct.c_double),
        ("temperature", ct.c_double),
        ("nc", ct.c_double),
        ("nv", ct.c_double),
        ("ec", ct.c_double),
        ("ev", ct.c_double),
        ("recombination", ct.c_int),
        ("rec_params", ct.c_double * 3),
        ("tau", ct.c_double * 3),
        ("flux_type", ct.c_int),
        ("phi_linearize", ct.c_int),
        ("adim", ct.c_double * 4),
        ("bc_type", (ct.c_int * 3) * 6),
        ("drift_jacobian", ct.c_int),
    ]


def use_library(path):
    """load the restatement from another build of hdg_oracle.cpp (the CPU
    baseline builds one for the host it runs on); Oracle objects created
    afterwards use it"""
    global _LIB
    _LIB = None
    return lib(path)


def lib(path=None):
    global _LIB
    if _LIB is None:
        path = path or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: build it with make -C {_HERE}")
        L = ct.CDLL(path)
        P = ct.POINTER
        d, i, ll, v = ct.c_double, ct.c_int, ct.c_longlong, ct.c_void_p
        dp, lp, ip = P(ct.c_double), P(ct.c_longlong), P(ct.c_int)
        sigs = {
            "orc_create": (i, [P(OrcProblem), P(v)]),
            "orc_destroy": (None, [v]),
            "orc_last_error": (ct.c_char_p, []),
            "orc_sizes": (None, [v, lp]),
            "orc_set_enabled": (i, [v, i, i, i]),
            "orc_cell_qpoints": (None, [v, dp]),
            "orc_face_qpoints": (None, [v, dp]),
            "orc_set_doping": (None, [v, dp]),
            "orc_set_bc_values": (None, [v, dp, dp]),
            "orc_set_state": (None, [v, dp, dp]),
            "orc_set_cell_face_tau": (None, [v, dp]),
            "orc_get_state": (None, [v, dp, dp]),
            "orc_local_system": (i, [v, i, i, dp, dp, dp, dp, dp, dp]),
            "orc_condensed_cell": (i, [v, i, i, dp, dp, lp, ip, lp, dp]),
            "orc_csr_pattern": (None, [v, lp, ip]),
            "orc_pass1": (d, [v, i, i, ll]),
            "orc_get_system": (None, [v, dp, dp]),
            "orc_n_dirichlet": (i, [v]),
            "orc_get_dirichlet": (None, [v, lp, dp]),
            "orc_pass2": (d, [v, i, i, dp, dp]),
            "orc_recombination": (None, [P(OrcProblem), d, d, d, dp]),
            "orc_terminal_current": (i, [v, i, dp]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _lp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_longlong))


def _ip(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int))


@dataclass
class NewtonResult:
    """NonlinearIterationResults (include/nonlinear_iteration_results.h:5-14)."""
    converged: bool
    iterations: int
    last_update_norm: float


class Oracle:
    """One mesh + problem in the CPU restatement."""

    def __init__(self, problem: OrcProblem):
        self.L = lib()
        self.ctx = ct.c_void_p()
        rc = self.L.orc_create(ct.byref(problem), ct.byref(self.ctx))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        self.problem = problem
        self._refresh()

    def __del__(self):
        try:
            if self.ctx:
                self.L.orc_destroy(self.ctx)
        except Exception:
            pass

    def _refresh(self):
        s = np.zeros(12, dtype=np.int64)
        self.L.orc_sizes(self.ctx, _lp(s))
        (self.ncells, self.nfaces, self.NC, self.Qc, self.Qf, self.nb, self.nf,
         self.ne, self.N, self.T, self.ntr, self.nnz) = (int(x) for x in s)
        self.dim = self.problem.dim

    def set_enabled(self, V, n, p):
        if self.L.orc_set_enabled(self.ctx, int(V), int(n), int(p)) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        self._refresh()

    def cell_qpoints(self):
        x = np.zeros((self.ncells, self.Qc, self.dim))
        self.L.orc_cell_qpoints(self.ctx, _dp(x))
        return x

    def face_qpoints(self):
        x = np.zeros((self.nfaces, self.Qf, self.dim))
        self.L.orc_face_qpoints(self.ctx, _dp(x))
        return x

    def set_doping(self, doping_qp):
        a = np.ascontiguousarray(doping_qp, dtype=np.float64)
        self.L.orc_set_doping(self.ctx, _dp(a))

    def set_bc_values(self, dirichlet=None, neumann=None):
        n = self.nfaces * 3 * self.Qf
        d = np.ascontiguousarray(dirichlet if dirichlet is not None else np.zeros(n), dtype=np.float64)
        m = np.ascontiguousarray(neumann if neumann is not None else np.zeros(n), dtype=np.float64)
        self.L.orc_set_bc_values(self.ctx, _dp(d), _dp(m))

    def set_cell_face_tau(self, tau):
        """CellFaceTauComputer values, physical, shape (ncells, 2*dim, 3)
        (cell, face, V/n/p); None returns to the problem's fixed tau"""
        if tau is None:
            self.L.orc_set_cell_face_tau(self.ctx, None)
            return
        a = np.ascontiguousarray(tau, dtype=np.float64)
        assert a.size == self.ncells * 2 * self.dim * 3
        self._tau_cf = a
        self.L.orc_set_cell_face_tau(self.ctx, _dp(a))

    def set_state(self, cell, trace):
        c = np.ascontiguousarray(cell, dtype=np.float64)
        t = np.ascontiguousarray(trace, dtype=np.float64)
        assert c.size == self.ncells * self.NC and t.size == self.nfaces * 3 * self.nf
        self.L.orc_set_state(self.ctx, _dp(c), _dp(t))

    def get_state(self):
        c = np.zeros(self.ncells * self.NC)
        t = np.zeros(self.nfaces * 3 * self.nf)
        self.L.orc_get_state(self.ctx, _dp(c), _dp(t))
        return c, t

    def local_system(self, cell, thermodyn=False):
        N, T = self.N, self.T
        A, CT, TC, TT = np.zeros((N, N)), np.zeros((N, T)), np.zeros((T, N)), np.zeros((T, T))
        c, t = np.zeros(N), np.zeros(T)
        rc = self.L.orc_local_system(self.ctx, cell, int(thermodyn), _dp(A), _dp(CT), _dp(TC),
                                     _dp(TT), _dp(c), _dp(t))
        assert rc == 0
        return A, CT, TC, TT, c, t

    def condensed_cell(self, cell, thermodyn=False):
        T = self.T
        S, g = np.zeros((T, T)), np.zeros(T)
        dofs = np.zeros(T, dtype=np.int64)
        nd = ct.c_int()
        di, dv = np.zeros(T, dtype=np.int64), np.zeros(T)
        rc = self.L.orc_condensed_cell(self.ctx, cell, int(thermodyn), _dp(S), _dp(g), _lp(dofs),
                                       ct.byref(nd), _lp(di), _dp(dv))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return S, g, dofs, di[: nd.value], dv[: nd.value]

    def csr_pattern(self):
        rp = np.zeros(self.ntr + 1, dtype=np.int64)
        ci = np.zeros(self.nnz, dtype=np.int32)
        self.L.orc_csr_pattern(self.ctx, _lp(rp), _ip(ci))
        return rp, ci

    def pass1(self, thermodyn=False, nthreads=1, limit=-1):
        t = self.L.orc_pass1(self.ctx, int(thermodyn), int(nthreads), int(limit))
        if t < 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return t

    def system(self):
        import scipy.sparse as sp
        rp, ci = self.csr_pattern()
        vals, rhs = np.zeros(self.nnz), np.zeros(self.ntr)
        self.L.orc_get_system(self.ctx, _dp(vals), _dp(rhs))
        n = self.L.orc_n_dirichlet(self.ctx)
        di, dv = np.zeros(n, dtype=np.int64), np.zeros(n)
        if n:
            self.L.orc_get_dirichlet(self.ctx, _lp(di), _dp(dv))
        return sp.csr_matrix((vals, ci, rp), shape=(self.ntr, self.ntr)), rhs, di, dv

    def pass2(self, update_trace_full, thermodyn=False, nthreads=1):
        u = np.ascontiguousarray(update_trace_full, dtype=np.float64)
        out = np.zeros(self.ncells * self.NC)
        t = self.L.orc_pass2(self.ctx, int(thermodyn), int(nthreads), _dp(u), _dp(out))
        if t < 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return out

    def terminal_current(self, boundary_id):
        """integrals of the numerical fluxes of n and p over the faces of
        boundary_id (dimensionless, outward normal)."""
        out = np.zeros(2)
        if self.L.orc_terminal_current(self.ctx, int(boundary_id), _dp(out)) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return out

    # ------------------------------------------------------------- helpers
    def restricted_to_full(self, enabled):
        """build_restricted_to_trace_dof_map (src/np_solver.cpp:1357-1432)."""
        nf = self.nf
        comps = [c for c in range(3) if enabled[c]]
        ne = len(comps)
        f = np.repeat(np.arange(self.nfaces), ne * nf)
        ci = np.tile(np.repeat(np.arange(ne), nf), self.nfaces)
        j = np.tile(np.arange(nf), self.nfaces * ne)
        comp = np.asarray(comps)[ci]
        return f * 3 * nf + comp * nf + j

    def run_newton(self, enabled=(True, True, True), thermodyn=False, abs_tol=1e-10,
                   rel_tol=1e-10, max_it=100, alpha=1.0, nthreads=1, solver=None,
                   log=None):
        """NPSolver::private_run (src/np_solver.cpp:1452-1561), including the
        L-infinity update criterion measured against the pre-update state."""
        import scipy.sparse.linalg as spla
        self.set_enabled(*enabled)
        r2f = self.restricted_to_full(enabled)
        cell, trace = self.get_state()
        converged, step, upd_cell_norm = False, 0, 0.0
        for step in range(1, max_it + 1):
            self.pass1(thermodyn, nthreads)
            A, rhs, di, dv = self.system()
            if solver is None:
                sol = spla.splu(A.tocsc()).solve(rhs)
            else:
                sol = solver(A, rhs)
            sol[di] = dv
            upd_trace = np.zeros(self.nfaces * 3 * self.nf)
            upd_trace[r2f] = sol
            upd_cell = self.pass2(upd_trace, thermodyn, nthreads)
            upd_cell_norm = float(np.max(np.abs(upd_cell))) if upd_cell.size else 0.0
            upd_norm = max(upd_cell_norm, float(np.max(np.abs(upd_trace))))
            cur_norm = max(float(np.max(np.abs(cell))), float(np.max(np.abs(trace))))
            cell = cell + alpha * upd_cell
            trace = trace + alpha * upd_trace
            self.set_state(cell, trace)
            if log is not None:
                log(step, upd_norm, cur_norm)
            if upd_norm < abs_tol or upd_norm < rel_tol * cur_norm:
                converged = True
                break
        return NewtonResult(converged, step, upd_cell_norm)


def recombination(problem: OrcProblem, n, p, rf):
    out = np.zeros(3)
    lib().orc_recombination(ct.byref(problem), float(n), float(p), float(rf), _dp(out))
    return out


def make_problem(dim, degree, n_cells, lower, upper, *, epsilon, mu_n, mu_p, temperature,
                 nc, nv, ec, ev, recombination=0, rec_params=(0, 0, 0), tau=(1, 1, 1),
                 flux_type=0, phi_linearize=False, adim=(1.0, 1.602176634e-19 / 1.380649e-23, 1.0, 1.0),
                 bc_type=None, drift_jacobian=0) -> OrcProblem:
    P = OrcProblem()
    P.dim, P.degree = dim, degree
    for a in range(3):
        P.n_cells[a] = n_cells[a] if a < dim else 1
        P.lower[a] = lower[a] if a < dim else 0.0
        P.upper[a] = upper[a] if a < dim else 1.0
    P.epsilon, P.mu_n, P.mu_p, P.temperature = epsilon, mu_n, mu_p, temperature
    P.nc, P.nv, P.ec, P.ev = nc, nv, ec, ev
    P.recombination = recombination
    for i in range(3):
        P.rec_params[i] = rec_params[i]
        P.tau[i] = tau[i]
    P.flux_type, P.phi_linearize = flux_type, int(phi_linearize)
    for i in range(4):
        P.adim[i] = adim[i]
    P.drift_jacobian = int(drift_jacobian)
    if bc_type is not None:
        for bid in range(6):
            for c in range(3):
                P.bc_type[bid][c] = bc_type[bid][c]
    return P
